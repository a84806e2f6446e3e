// jdem_internal.h -- context, scratch arena and kernel-argument structs shared
// by the translation units of libjdem_hip.so.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "jdem.h"
#include "jdem_device.h"

namespace jdem {

constexpr int kBlock = 256;
// Workgroups of a launch whose size is known only on the device (gated
// batch rebuilds): its kernels walk the work in turn, so an idle launch
// dispatches this many workgroups instead of one per tile.
constexpr int kCappedGrid = 2048;

// Grid parameters on the device (jdem_grid; jdem_grid_f64 for double).
template <class Real = float>
struct GridPT {
  int dims[3];
  int strides[3];
  Real cell[3];
  int sr;
  int periodic;
  int64_t n_cells;
};
using GridP = GridPT<float>;

// One grow-only device allocation of the context.
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace jdem

struct jdem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // scratch, grown on demand and reused across calls
  jdem::Buf pos2, vel2, ang2;          // ping-pong kinematics
  jdem::Buf keys[2], vals[2], hist;    // radix sort
  jdem::Buf ccur;                      // counting sort: per-cell cursors
  jdem::Buf cstart, cend, spos;        // cell tables, sorted (x, r)
  jdem::Buf nl_new, xi_new, live_new;  // neighbour-list rebuild
  jdem::Buf flags;                     // small device words
  jdem::Buf red;                       // reduction partials
  jdem::Buf nl_copy, xi_copy, live_copy;
  // resident NeighborList working set (jdem_resident.hip)
  jdem::Buf wP[2], wV[2], wW[2], wQ[2], wI[2], wF, wT, wOld, wOrig[2], wWpos, wSrc, wInv, wKeys, fireV, fireDev, segp, segid, segflag, wNl[2], wXi[2], wLive[2], wXiRank;
  int32_t* host_flags = nullptr;       // pinned host mirror of flags
  // HIP-event timing (jdem_ctx_set_timing)
  struct TimedSpan {
    hipEvent_t start, stop;
    int kind;  // 0 step kernel, 1 idle step kernel, 2 rebuild
    int64_t step;
  };
  bool timing = false;
  std::vector<hipEvent_t> event_pool;
  std::vector<TimedSpan> spans;
  // Device buffers not in use: released working sets (jdem_resident) and
  // buffers handed over to them leave empty scratch slots that ensure()
  // refills from here before it allocates.
  std::vector<jdem::Buf> pool;
  std::mutex pool_mu;
  // private stream on which captured step graphs are replayed (default priority)
  hipStream_t gstream = nullptr;
  hipEvent_t gevent[2] = {nullptr, nullptr};
  // device copy of the last routed law's terms and species offsets, and the
  // host bytes it was made from
  jdem::Buf route;
  std::vector<char> route_host;
  // per-cell counts / starts of the batch rebuild's counting sort
  jdem::Buf ccnt;
  // clumps: members per clump (CSR), lab positions, levers, member totals
  jdem::Buf cl_start, cl_members, plab, ppr, ftot, ttot, crefl;
  // the precision form's Verlet list and its per-sphere build state
  jdem::Buf pvl, pvaux;
  // every scratch buffer of the context, for accounting and release
  std::vector<jdem::Buf*> bufs();
  size_t scratch_bytes();
};

// The NeighborList working set kept between calls (include/jdem.h): the
// work order and the buffers of the step kernel's layouts, in work order.
struct jdem_resident {
  jdem_ctx* ctx = nullptr;
  int64_t n = 0;
  int M = 0;
  int W64 = 1;
  int dim = 3;
  bool hist = false;
  jdem::Buf orig;  // int32 (N): original index of work slot b
  jdem::Buf old4;  // float4 (N): positions at the last build
  jdem::Buf nl;    // int32 slot-major (M, N): rows end at their first -1
  jdem::Buf live;  // uint64 (W64, N): slots that carry a spring
  jdem::Buf xi;    // float4 by rank (M, N): k-th live spring of slot b
};

namespace jdem {

// Kernel-side parameters of the ABI's domain and grid structs.
DomP dom_of(const jdem_domain& d);
GridP grid_of(const jdem_grid& g);
DomPT<double> dom_of(const jdem_domain_f64& d);
GridPT<double> grid_of(const jdem_grid_f64& g);

// Ensure b holds at least `bytes`; returns false on allocation failure.
// An empty b is first refilled from the context's pool.
bool ensure(jdem_ctx* ctx, Buf& b, size_t bytes);
// Give b's memory to the context's pool and leave b empty.
void pool_give(jdem_ctx* ctx, Buf& b);
// Free every pooled buffer (after the stream has drained).
void pool_trim(jdem_ctx* ctx);

// Kernel arguments of the fused step / force / energy kernels.
template <class Real = float>
struct StepArgsT {
  int64_t n;
  int bond_w;
  int isb;
  const Real* pos_in;
  const Real* vel_in;
  const Real* ang_in;
  Real* pos_out;
  Real* vel_out;
  Real* ang_out;
  Real* q;
  Real* force;
  Real* torque;
  const Real* rad;
  const Real* mass;
  const Real* inertia;
  const int32_t* mat_id;
  const int32_t* bond_id;
  const uint8_t* fixed;
  const Real* ext_f;
  const Real* ext_fc;
  const Real* ext_t;
  Real g[3];
  int has_fm;
  LawPT<Real> law;
  DomPT<Real> dom;
  Real dt;
  int lin;
  int rot;
  // cell-list source
  const int32_t* perm;
  const int32_t* skeys;
  const int32_t* cstart;
  const int32_t* cend;
  const float4* spos;  // float only: sorted (x, r)
  GridPT<Real> grid;
  // neighbour-list source
  const int32_t* nl;
  int M;
  Real* hist_xi;
  uint32_t* hist_live;
  int W32;
  // phases
  int do_force;
  int force_zero;  // collider "": contact force and torque are zero
  int do_after;
  int do_before;
  int write_kin;
  // neighbour-list displacement check (neighbor_list.py:78-85)
  const Real* old_pos;
  Real trigger;
  int check_disp;
  const int32_t* flag_rd;
  int32_t* flag_wr;
  int32_t step_idx;
  // energy / virial outputs (per particle, float64)
  double* e_out;
  // clumps (null for spheres): ids (the pair mask), R(q) pos_p per member
  // (the torque lever), the members of each clump (CSR: cl_start (N + 1),
  // cl_members (N) in ascending index), body offsets, and the per-member
  // force / torque totals of ForceManager before the clump sum
  const int32_t* clump;
  Real* ppr;
  const int32_t* cl_start;
  const int32_t* cl_members;
  const Real* pos_p;
  Real* plab;
  Real* ftot;
  Real* ttot;
  Real* rscr;  // reflective domain with clumps: per-member scratch (kReflW words)
};
using StepArgs = StepArgsT<float>;
// The float64 form of the plain step (jdem_system_step_f64 and friends).
using StepArgs64 = StepArgsT<double>;
using DomP64 = DomPT<double>;
using GridP64 = GridPT<double>;
using LawP64 = LawPT<double>;

enum Src { SRC_NAIVE = 1, SRC_CELLS = 2, SRC_NLIST = 3 };

// launchers (jdem_step.hip)
hipError_t launch_step(const StepArgs& a, int dim, int law, int src, hipStream_t s);
hipError_t launch_energy(const StepArgs& a, int dim, int law, int src, hipStream_t s);
hipError_t launch_virial(const StepArgs& a, int dim, int law, int src, hipStream_t s);
// float64 instances: every collider source, laws without history
hipError_t launch_step(const StepArgs64& a, int dim, int law, int src, hipStream_t s);
hipError_t launch_energy(const StepArgs64& a, int dim, int law, int src, hipStream_t s);
hipError_t launch_kinetic(const StepArgs64& a, int dim, hipStream_t s);
// clumps: lab positions and levers; ForceManager totals; the sum over each
// clump broadcast to its members (both scalar types)
hipError_t launch_clump_lab(const StepArgs& a, int dim, hipStream_t s);
hipError_t launch_clump_lab(const StepArgs64& a, int dim, hipStream_t s);
hipError_t launch_clump_forces(const StepArgs& a, int dim, hipStream_t s);
hipError_t launch_clump_forces(const StepArgs64& a, int dim, hipStream_t s);
// ReflectDomain.apply for clumps (four launches; scratch a.rscr of
// refl_scratch_words(dim) words per member)
hipError_t launch_clump_reflect(const StepArgs& a, int dim, hipStream_t s);
hipError_t launch_clump_reflect(const StepArgs64& a, int dim, hipStream_t s);
inline int refl_scratch_words(int dim) { return 8 * dim + 2 + (dim == 3 ? 3 : 1); }

// launchers (jdem_cells.hip)
hipError_t launch_hash(const float* pos, int64_t n, int dim, const DomP& dom, const GridP& g, uint32_t* keys,
                       hipStream_t s);
hipError_t launch_hash(const double* pos, int64_t n, int dim, const DomP64& dom, const GridP64& g, uint32_t* keys,
                       hipStream_t s);
// Buffers of a sort: key and value ping-pong arrays of n entries, a
// histogram / scan-sum array and (counting sort) per-cell cursors.
struct SortBufs {
  uint32_t* k0;
  uint32_t* k1;
  int32_t* v0;
  int32_t* v1;
  uint32_t* hist;
  int32_t* cur;
};
// uint32 words of SortBufs::hist that a radix sort of n keys uses
size_t radix_hist_words(int64_t n);
// uint32 words of SortBufs::hist that a counting sort over n_cells uses
size_t cell_sort_sum_words(int64_t n_cells);
// The sorts below on caller-given buffers (no allocation).
// With n_dev the key count is read on the device (n bounds it and sizes the
// launches); with gate every kernel returns at once while *gate == 0.
hipError_t radix_sort_bufs(const SortBufs& sb, const uint32_t* keys_in, int64_t n, int nbits, uint32_t** keys_out,
                           int32_t** vals_out, hipStream_t s, const int64_t* n_dev = nullptr,
                           const int32_t* gate = nullptr);
hipError_t cell_sort_bufs(const SortBufs& sb, const uint32_t* keys_in, int64_t n, int64_t n_cells, uint32_t** keys_out,
                          int32_t** vals_out, int32_t* starts, hipStream_t s, const int32_t* gate = nullptr);
// Sort (keys, iota) by key; results in (keys_out, vals_out).  nbits = key bits.
hipError_t radix_sort_pairs(jdem_ctx* ctx, const uint32_t* keys_in, int64_t n, int nbits, uint32_t** keys_out,
                            int32_t** vals_out, hipStream_t s, const int64_t* n_dev = nullptr,
                            const int32_t* gate = nullptr);
// The same result as radix_sort_pairs for keys < n_cells, by counting:
// per-cell counts, their exclusive scan (written to starts[0..n_cells],
// starts[n_cells] = n), placement through per-cell cursors, then each
// element's rank among its cell's ids.  Cost grows with n + n_cells.
// With gate, every kernel of the sort returns at once while *gate == 0 (a
// device word), leaving the previous sort's outputs in place.
hipError_t cell_sort_pairs(jdem_ctx* ctx, const uint32_t* keys_in, int64_t n, int64_t n_cells, uint32_t** keys_out,
                           int32_t** vals_out, int32_t* starts, hipStream_t s, const int32_t* gate = nullptr);
// The counting sort with both counts on the device: the first *n_dev of
// the keys (n bounds it), every key below *c_dev (n_cells bounds it).  The
// launches are sized by the bounds and capped (kCappedGrid workgroups that
// walk the work in turn), so a rebuild of a few systems of a large batch
// costs what its particles and cells need.  Results as cell_sort_pairs; the
// cell starts (compact, *c_dev + 1 entries) go to the context's scratch.
hipError_t cell_sort_dev(jdem_ctx* ctx, const uint32_t* keys_in, int64_t n, int64_t n_cells, const int64_t* n_dev,
                         const int64_t* c_dev, uint32_t** keys_out, int32_t** vals_out, hipStream_t s,
                         const int32_t* gate);
// cend[c] = starts[c + 1] for c < n_cells
hipError_t launch_starts_to_ends(const int32_t* starts, int64_t n_cells, int32_t* cend, hipStream_t s);
// key bits of a grid of n_cells cells (at least 1)
inline int nbits_for_cells(int64_t n_cells) {
  int64_t m = n_cells > 1 ? n_cells - 1 : 1;
  int b = 0;
  while (m > 0) {
    ++b;
    m >>= 1;
  }
  return b;
}
// Device address of the fixed scratch array of jdem_cell_build_noalloc,
// looked up once per device (jdem_noalloc.hip).
int noalloc_init(int device);
// whether cell_sort_pairs is the cheaper sort for n keys below n_cells
inline bool use_cell_sort(int64_t n, int64_t n_cells) { return n > 0 && n_cells <= 4 * n + 65536; }
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
hipError_t launch_cell_ranges(const uint32_t* skeys, int64_t n, int32_t* cstart, int32_t* cend, hipStream_t s);
hipError_t launch_gather_spos(const int32_t* perm, const float* pos, const float* rad, int64_t n, int dim,
                              float4* spos, hipStream_t s);
template <class Real = float>
struct NLBuildT {
  int64_t n;
  int dim;
  int M;
  Real cutoff_sq;
  int isb;
  int bond_w;
  const int32_t* bond_id;
  const Real* pos;
  const int32_t* perm;
  const int32_t* skeys;
  const int32_t* cstart;
  const int32_t* cend;
  const float4* spos;  // float only: sorted (x, r)
  GridPT<Real> grid;
  DomPT<Real> dom;
  int32_t* nl;
  int32_t* overflow;
  const int32_t* clump;  // clump ids (pairs of one clump are not listed), or null
};
using NLBuild = NLBuildT<float>;
hipError_t launch_nlist_cells(const NLBuild& b, hipStream_t s);
hipError_t launch_nlist_naive(const NLBuild& b, hipStream_t s);
hipError_t launch_nlist_cells(const NLBuildT<double>& b, hipStream_t s);
hipError_t launch_nlist_naive(const NLBuildT<double>& b, hipStream_t s);
hipError_t launch_remap_hist(int64_t n, int M, int W32, int dim, const int32_t* old_nl, const uint32_t* old_live,
                             const float* old_xi, const int32_t* new_nl, uint32_t* new_live, float* new_xi,
                             hipStream_t s);
// pos_p (clumps, or null): the bounding radius is |pos_p| + rad
hipError_t launch_box_minmax(const float* pos, const float* rad, int64_t n, int dim, float* part, int nparts,
                             hipStream_t s, const float* pos_p = nullptr);
hipError_t launch_box_minmax(const double* pos, const double* rad, int64_t n, int dim, double* part, int nparts,
                             hipStream_t s, const double* pos_p = nullptr);
hipError_t launch_sum_f64(const double* in, int64_t n, int stride, int ncomp, double* partial, double* out,
                          int nparts, hipStream_t s);
hipError_t launch_kinetic(const StepArgs& a, int dim, hipStream_t s);
hipError_t launch_stage(const StepArgs& a, int dim, int stage, hipStream_t s);
hipError_t launch_copy_f32(float* dst, const float* src, int64_t n, hipStream_t s);


// ---- resident NeighborList engine (jdem_resident.hip) ----------------------
struct InOut {
  int64_t n;
  int dim;
  int M;
  int W64;  // live words per particle
  // caller arrays (original order)
  float* pos;
  float* vel;
  float* ang;
  float* q;
  float* force;
  float* torque;
  const float* rad;
  const float* mass;
  const float* inertia;
  const int32_t* mat_id;
  const int32_t* species;  // routed laws with species, else null
  const uint8_t* fixed;
  int32_t* nl_user;      // (N, M)
  float* old_user;       // (N, D)
  float* xi_user;        // (N, M, D) or null
  uint32_t* live_user;   // (N, ceil(M/32)) or null
  // work arrays
  const int32_t* orig;  // work -> original
  const int32_t* wpos;  // original -> work
  float4* P4;
  float4* V4;
  float4* W4;
  float4* Q4;
  float4* I4;
  float4* F4;
  float4* T4;
  float4* old4;
  int32_t* nl;
  float4* xi;
  uint64_t* live;
  int old_keep;  // entry: old4 already holds the build positions
  // entry: cleared when some sphere is fixed or its inertia differs from
  // (0.4 m) r^2 in any component (3D; set to 1 before the entry)
  int32_t* iso_flag;
};

// Parameters of one system of a segmented batch (jdem_batch_step).  The
// systems' particles occupy consecutive ranges of the work set in both the
// caller's order and the cell order, because each system owns a consecutive
// range of work cell keys starting at cell_base.
struct SegP {
  DomP dom;
  GridP grid;
  uint32_t cell_base;
  float dt;
  float cutoff_sq;  // (cutoff + skin)^2 of the list build
  float trigger;    // rebuild threshold on |x - x_build|^2
};

struct NLW {
  int64_t n;
  int M;
  float cutoff_sq;
  int isb;
  int bond_w;
  const int32_t* bond_id;  // original indexing
  const int32_t* orig;
  const float4* P4;
  const uint32_t* skeys;  // sorted keys = cell of work slot
  const int32_t* starts;  // (n_cells + 1) first slot of each cell
  GridP grid;
  DomP dom;
  int32_t* nl;
  int32_t* overflow;
  // history carried over from the previous list (null live: no history)
  int W64;
  int remap;                // 0: start with empty history, 1: remap old
  const int32_t* src;       // new work slot -> old work slot
  const int32_t* inv;       // old work slot -> new work slot
  const int32_t* old_nl;
  const float4* old_xi;
  // old springs per wavefront (the step kernel's layout), their offsets in
  // old_W[.].w; null: old_xi is by rank
  const float4* old_W;
  // new springs per wavefront with their offsets written to new_W[.].w
  // (one wave-wide scan of the new counts); null: xi is by rank
  float4* new_W;
  const uint64_t* old_live;
  float4* xi;
  uint64_t* live;
  // segmented batch: system of each work slot, the systems' parameters and
  // what this rebuild does to each system's rows (SEG_KEEP / SEG_FRESH /
  // SEG_REMAP)
  const int32_t* seg;
  const SegP* segp;
  const int32_t* seg_mode;
  // batches rebuild in place: the old rows, springs and live words of the
  // particle now at b are at compact index cidx[b] (stride ostride)
  const int32_t* cidx;
  int64_t ostride;
  const int64_t* ccref;  // compact key of system s's first cell
  const int32_t* gate;   // when set, the kernel returns at once while *gate == 0
  // with wpref (flagged systems' wavefront prefix), the launch walks only the
  // wavefronts that hold a flagged system's slots
  const int64_t* wpref;
  const int64_t* offsets;
  int n_seg;
};
enum { SEG_KEEP = 0, SEG_FRESH = 1, SEG_REMAP = 2 };

// In-place rebuild of the flagged systems of a batch (jdem_resident.hip).
struct RB {
  int64_t n, fn;  // particles; flagged particles (-1: read on the device, cpref[n_seg])
  int n_seg;
  int dim;
  const int32_t* gate;  // when set, every kernel returns at once while *gate == 0
  int M, W64;
  const int32_t* seg;
  const SegP* segp;
  const int32_t* seg_mode;
  const int64_t* offsets;
  const int64_t* cpref;
  const int64_t* ccref;
  // work set (updated in place)
  float4 *P, *V, *W, *Q, *I, *old4;
  int32_t* orig;
  const int32_t* nl;
  const uint64_t* live;
  const float4* xi;
  // compact sort input / output
  uint32_t* keys_c;
  int32_t* slot_c;
  const int32_t* perm;
  // staging, compact index
  float4 *tP, *tV, *tW, *tQ, *tI;
  int32_t* torig;
  int32_t* cs;
  int32_t* tnl;
  uint64_t* tlive;
  float4* txi;
  // maps
  int32_t* inv;
  int32_t* cidx;
};
hipError_t ws_launch_seg_prefix(int n_seg, const int32_t* seg_mode, const int64_t* offsets, const SegP* segp,
                               int64_t* cpref, int64_t* ccref, int64_t* wpref, hipStream_t s);
hipError_t ws_launch_hash_flagged(const RB& r, int dim, hipStream_t s);
hipError_t ws_launch_place(const RB& r, const uint32_t* sk, int32_t* starts, hipStream_t s);

// Workgroups are dispatched round-robin over the 8 XCDs of the chip, each
// with its own L2: workgroup w runs on XCD w % 8.  xcd_tile maps them so
// that XCD x processes one contiguous run of tiles; in the cell-sorted work
// order, spatial neighbours are then mostly served by the same L2.
__device__ inline unsigned xcd_tile(unsigned w, unsigned nw) {
  const unsigned q = nw / 8u, r = nw % 8u, x = w % 8u, l = w / 8u;
  return (x < r ? x * (q + 1u) : r * (q + 1u) + (x - r) * q) + l;
}

struct WS {
  int64_t n;
  int M;
  int W64;
  // 3D: device word, nonzero when every sphere is free with inertia (0.4 m) r^2
  // in each component and one material: the kernel computes the inertia
  // record instead of reading it
  const int32_t* iso;
  LawP law;
  DomP dom;
  float dt;
  int lin, rot;
  float g[3];
  int has_fm;
  const float* ext_f;   // original order
  const float* ext_fc;
  const float* ext_t;
  const int32_t* orig;
  const float4* P;
  const float4* V;
  const float4* W;
  float4* P2;
  float4* V2;
  float4* W2;
  float4* Q;
  const float4* I;
  float4* F;
  float4* T;
  const float4* old4;
  const int32_t* nl;
  const float4* xi_rd;  // springs by rank, read
  float4* xi_wr;        // springs by rank, written (another buffer)
  uint64_t* live;
  int do_force, do_after, do_before, store_ft;
  float trigger;
  int check_disp;
  const int32_t* flag_rd;
  int32_t* flag_wr;
  int32_t step_idx;
  // segmented batch (null for one system); seg_flag[s] = 1 when system s
  // needs a rebuild
  const int32_t* seg;
  const SegP* segp;
  int32_t* seg_flag;
  int block;  // workgroup size of a batch's step kernel (256 or 128)
};

hipError_t ws_launch_in(const InOut& io, bool with_nl, hipStream_t s);
hipError_t ws_launch_out(const InOut& io, bool with_nl, hipStream_t s);
// A working set's list, build positions and springs in the caller's layouts
// (io.nl_user / old_user / xi_user / live_user; each may be null).
hipError_t ws_launch_export(const InOut& io, hipStream_t s);
// dst[s*n + b] = src[s*n + b] for the slots of row b up to its first -1
hipError_t ws_launch_copy_rows(const int32_t* src, int32_t* dst, int64_t n, int M, hipStream_t s);
// Tangential springs of the step kernel, per wavefront: the 64 work slots
// of wave w own the region [64 w M, 64 (w + 1) M) of the spring buffer, and
// slot b's springs (in slot order of its live bits) start at offset W[b].w
// (an int in the float bits) of that region.  The list rebuild, the entry
// and the exit keep springs by rank instead ([k N + b], k-th live slot).
// pack: by rank -> per wave (offsets = prefix of the live counts);
// unpack: per wave -> by rank.  With seg_mode, only wavefronts holding a
// slot of a system that is rebuilt (mode != SEG_KEEP) are converted; with
// wpref (the flagged systems' wavefront prefix of k_seg_prefix) the launch
// walks those wavefronts only.
hipError_t ws_launch_pack_springs(int64_t n, int M, int W64, const float4* rank, const uint64_t* live, float4* wave,
                                  float4* W, const int32_t* seg, const int32_t* seg_mode, hipStream_t s,
                                  const int32_t* gate = nullptr, const int64_t* wpref = nullptr,
                                  const int64_t* offsets = nullptr, int n_seg = 0);
hipError_t ws_launch_unpack_springs(int64_t n, int M, int W64, const float4* wave, const uint64_t* live,
                                    const float4* W, float4* rank, const int32_t* seg, const int32_t* seg_mode,
                                    hipStream_t s, const int32_t* gate = nullptr, const int64_t* wpref = nullptr,
                                    const int64_t* offsets = nullptr, int n_seg = 0);
hipError_t ws_launch_inverse(const int32_t* orig, int64_t n, int32_t* wpos, hipStream_t s);
hipError_t ws_launch_iota(int32_t* p, int64_t n, hipStream_t s);
hipError_t ws_launch_copy4(float4* d, const float4* src, int64_t n, hipStream_t s);
hipError_t ws_launch_hash(const float4* P4, const int32_t* orig, int64_t n, int dim, const DomP& dom, const GridP& g,
                          uint32_t* keys, int32_t* wpos, hipStream_t s);
// seg[b] = system of particle b for the ranges offsets[0..n_seg]
hipError_t ws_launch_seg_ids(const int64_t* offsets, int n_seg, int64_t n, int32_t* seg, hipStream_t s);
hipError_t ws_launch_permute(int64_t n, const int32_t* new_orig, const int32_t* wpos, const float4* P,
                             const float4* V, const float4* W, const float4* Q, const float4* I, float4* P2,
                             float4* V2, float4* W2, float4* Q2, float4* I2, int32_t* orig2, int32_t* src,
                             int32_t* inv, hipStream_t s, float4* old2 = nullptr, int dim = 3);
// batch rebuilds: per-system mode from the flags (and bump the build counts),
// and the flags of never-built systems
// With gate set, *events counts the calls that find *gate != 0 (rebuilds).
hipError_t ws_launch_seg_modes(int n_seg, int32_t* seg_flag, int32_t* n_build, int32_t* seg_mode, hipStream_t s,
                               const int32_t* gate = nullptr, int32_t* events = nullptr);
hipError_t ws_launch_seg_init_flags(int n_seg, const int32_t* n_build, int32_t* seg_flag, hipStream_t s,
                                   int32_t* gate = nullptr);
// *gate = 0 (the end of a gated rebuild); *word = 0 when *gate is set
hipError_t ws_launch_gate_end(int32_t* gate, hipStream_t s);
hipError_t ws_launch_gated_zero(const int32_t* gate, int32_t* word, hipStream_t s);

hipError_t ws_launch_cell_starts(const uint32_t* skeys, int64_t n, int64_t n_cells, int32_t* starts, hipStream_t s);
hipError_t ws_launch_nlist(const NLW& a, int dim, bool naive, hipStream_t s);
hipError_t ws_launch_step(const WS& a, int dim, int law, hipStream_t s);
hipError_t ws_launch_box(const float4* P, int64_t n, float* part, int nparts, hipStream_t s);

// ---- FIRE minimiser (jdem_fire.hip) -----------------------------------------
struct FireP {
  float dt0, alpha_init, f_inc, f_dec, f_alpha;
  float dt_max, dt_min;  // dt0 * dt_{max,min}_scale, rounded once to float
  int n_min, n_bad_max;
};
// Optimiser scalars, kept on the device between the kernels of an iteration.
struct FireDev {
  float dt, alpha;
  int n_good, n_bad;
  float new_dt, new_alpha, dt_rev, vs;
  double power;
  unsigned max_grad_bits;  // max |grad| component as float bits
  int pad;
};
struct FireArgs {
  int64_t n;
  int dim;
  FireP p;
  FireDev* st;
  const float* force;
  const float* torque;
  const uint8_t* fixed;
  float* vel_pos;
  float* vel_rot;
  float* pos;
  float* q;
  double* part;
  int nparts;
  // potential energy of gravity (force_manager.py:450-459)
  float g[3];
  const float* mass;
  double* e_out;
  // when set, every kernel returns at once once *stop is set (device-side
  // stopping tests of the precision form)
  const int32_t* stop;
};
// Stopping state of a FIRE loop decided on the device (routines.py:284-314
// tests plus a stall test): k_prec_check runs before every update.
struct PrecStop {
  double pe, prev, best_g, best_e, gmax;
  long long best_at, steps;
  int32_t stop;
  int32_t first;
};
struct PrecStopP {
  double pe_tol, pe_diff_tol, force_tol, dN;
  long long max_steps, stall_steps;
};
hipError_t launch_fire_power(const FireArgs& a, hipStream_t s);
hipError_t launch_fire_decide(const FireArgs& a, hipStream_t s);
hipError_t launch_fire_move(const FireArgs& a, hipStream_t s);
hipError_t launch_fire_commit(const FireArgs& a, hipStream_t s);
// max |force| / |torque| component into st->max_grad_bits; gravity energy
// added into e_out (per particle) when g != 0
hipError_t launch_fire_grad_max(const FireArgs& a, hipStream_t s);
hipError_t launch_fire_gravity_energy(const FireArgs& a, hipStream_t s);

// ---- precision form: integer cell + float32 offset (jdem_fire.hip) ----------
struct PrecArgs {
  int64_t n;
  int dim;
  int dims[3];   // cells per axis (periodic)
  float h[3];    // cell edges: h + hlo as an unevaluated float pair
  float hlo[3];
  int32_t* cidx; // (N, D) cell of each sphere
  float* off;    // (N, D) position inside the cell, [0, h)
  const float* rad;
  float k;
  uint32_t* keys;        // cell key of each sphere
  const int32_t* perm;   // cell-sorted position -> sphere
  const int32_t* starts; // first sorted position of each cell (n_cells + 1)
  float4* sp;            // cell-sorted (offset, radius), gathered per evaluation
  float* force;          // (N, D)
  double* e_out;         // (N) half the pair energies of each sphere
  double* w_out;         // (N, D*D) half the pair virials, or null
  int32_t* moved;        // set when a sphere changed cell
  unsigned* gmax_bits;   // max|F| word of the FIRE state, cleared by the gather
  int gated;             // 1: the key kernel runs only while *moved, and the gather clears *moved
  const int32_t* stop;   // when set, every kernel returns at once once *stop is set
  // Verlet list (vl set): *moved means "rebuild the list", raised when a
  // sphere has moved skin / 2 since the build; the fold of the offsets, the
  // sort and the list build run only then, and the list force clears *moved
  int32_t* vl;           // (N, vm) by sorted slot: partner slot | cell difference << 24, -1 ends a row
  int vm;
  int vsr;               // stencil half-width of the list build (1 or 2)
  float vcut;            // list reach beyond r_i + r_j (the skin)
  float vtrig;           // (skin / 2)^2
  float* boff;           // (N, D) offsets at the build
  int32_t* vovf;         // set when a row would exceed vm
};
// the stopping tests of one FIRE iteration from the last evaluation's float64
// energy sum and max|F| word; stops, or counts the update that follows
hipError_t launch_prec_check(PrecStop* ps, const double* pe_sum, const unsigned* gmax_bits, const PrecStopP& p,
                             hipStream_t s);
hipError_t launch_prec_keys(const PrecArgs& a, hipStream_t s);
hipError_t launch_prec_gather(const PrecArgs& a, hipStream_t s);
hipError_t launch_prec_force(const PrecArgs& a, hipStream_t s);
hipError_t launch_prec_vbuild(const PrecArgs& a, hipStream_t s);
hipError_t launch_prec_vforce(const PrecArgs& a, hipStream_t s);
hipError_t launch_prec_move(const FireArgs& f, const PrecArgs& a, hipStream_t s);

inline unsigned nblocks(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

}  // namespace jdem
