// jdem_fire.hip -- the FIRE energy minimiser on the GPU
// (minimizers/optimizers.py:150-339, minimizers/routines.py:150-383).
//
// One iteration is four launches around the force evaluation:
//   k_fire_power   per particle: F = -grad (masked by ~fixed), v_old =
//                  v + F dt/2; block partial sums of F . v_old in float64
//   k_fire_decide  one block: the power, then the downhill / uphill update of
//                  dt, alpha, N_good, N_bad (optimizers.py:250-296)
//   k_fire_move    per particle: velocity mixing, the parameter update and
//                  the new orientation (routines.py:341-360, 37-65)
//   k_fire_commit  dt, alpha <- the values just decided
// The optimiser scalars stay on the device (FireDev); the host reads only the
// energy and the largest gradient component once per iteration.  Every
// float32 expression follows the reference's operation order, so positions
// match the float32 oracle bit for bit while the power decisions agree.
#include <hip/hip_runtime.h>

#include "jdem_internal.h"

namespace jdem {

namespace {

// optax.safe_norm(x, 1e-16, axis=-1): the row norm, or 1e-16 where the norm
// is not above it
template <int A>
JD float safe_norm(const float* x) {
  float n2;
  if (A == 3)
    n2 = (x[0] * x[0] + x[1] * x[1]) + x[2] * x[2];
  else if (A == 2)
    n2 = x[0] * x[0] + x[1] * x[1];
  else
    n2 = x[0] * x[0];
  const float n = sqrtf(n2);
  return n <= 1e-16f ? 1e-16f : n;
}

// Quaternion.from_rotvec (quaternion.py:99-127)
JD Quat q_from_rotvec(Vec<3> r) {
  const float n2 = dot(r, r);
  const float safe = fmaxf(n2, 1e-16f);
  const float theta = sqrtf(safe);
  const float half = 0.5f * theta;
  const float ch = cosf(half);
  const float sc = sinf(half) / theta;
  return {ch, r.x * sc, r.y * sc, r.z * sc};
}

// One parameter leaf of one particle (A components): F, v_old and the
// velocity-mixing update of optimizers.py:245-318.
template <int A>
struct Leaf {
  float f[A], v_old[A];
  JD void load(const float* F, const float* V, int64_t i, float m, float dt) {
#pragma unroll
    for (int c = 0; c < A; ++c) {
      f[c] = -((-F[A * i + c]) * m);  // F = -(grad * mask), grad = -force
      v_old[c] = V[A * i + c] + (f[c] * dt) / 2.0f;
    }
  }
  JD double power() const {
    double p = 0.0;
#pragma unroll
    for (int c = 0; c < A; ++c) p += (double)f[c] * (double)v_old[c];
    return p;
  }
  // returns the update; writes the new velocity
  JD void mix(const FireDev& s, float* V, int64_t i, float* upd) const {
    float vh[A];
#pragma unroll
    for (int c = 0; c < A; ++c) vh[c] = v_old[c] * s.vs + (f[c] * s.new_dt) / 2.0f;
    const float vn = safe_norm<A>(vh);
    const float fn = safe_norm<A>(f);
    const float mixr = fn > 1e-16f ? (vn / fn) * s.new_alpha : 0.0f;
#pragma unroll
    for (int c = 0; c < A; ++c) {
      vh[c] = (vh[c] * (1.0f - s.new_alpha) + f[c] * mixr) * s.vs;
      upd[c] = (v_old[c] * s.dt_rev) / 2.0f + (vh[c] * s.new_dt) / 2.0f;
      V[A * i + c] = vh[c];
    }
  }
};

template <int D>
__global__ void __launch_bounds__(kBlock) k_fire_power(FireArgs a) {
  if (a.stop && *a.stop) return;
  constexpr int A = D == 3 ? 3 : 1;
  __shared__ double sh[kBlock];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double p = 0.0;
  if (i < a.n) {
    const float m = a.fixed[i] ? 0.0f : 1.0f;
    const float dt = a.st->dt;
    Leaf<D> lp;
    Leaf<A> lr;
    lp.load(a.force, a.vel_pos, i, m, dt);
    lr.load(a.torque, a.vel_rot, i, m, dt);
    p = lp.power() + lr.power();
  }
  sh[threadIdx.x] = p;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_fire_decide(FireArgs a) {
  if (a.stop && *a.stop) return;
  __shared__ double sh[kBlock];
  double acc = 0.0;
  for (int b = threadIdx.x; b < a.nparts; b += kBlock) acc += a.part[b];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  FireDev s = *a.st;
  const double power = sh[0];
  const float cand_inc = fminf(s.dt * a.p.f_inc, a.p.dt_max);
  const float cand_dec = fmaxf(s.dt * a.p.f_dec, a.p.dt_min);
  if (power > 0.0) {  // downhill (optimizers.py:278-292)
    const int ng = s.n_good + 1;
    s.new_dt = ng > a.p.n_min ? cand_inc : s.dt;
    s.new_alpha = ng > a.p.n_min ? s.alpha * a.p.f_alpha : s.alpha;
    s.n_good = ng;
    s.n_bad = 0;
    s.dt_rev = 0.0f;
    s.vs = 1.0f;
  } else {  // uphill (optimizers.py:258-276)
    int nb = s.n_bad + 1;
    const bool exceeded = nb > a.p.n_bad_max;
    s.new_dt = exceeded ? a.p.dt0 : cand_dec;
    s.n_bad = exceeded ? 0 : nb;
    s.new_alpha = a.p.alpha_init;
    s.n_good = 0;
    s.dt_rev = -s.new_dt;
    s.vs = 0.0f;
  }
  s.power = power;
  *a.st = s;
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_fire_move(FireArgs a) {
  constexpr int A = D == 3 ? 3 : 1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const FireDev s = *a.st;
  const bool fixed = a.fixed[i] != 0;
  const float m = fixed ? 0.0f : 1.0f;
  Leaf<D> lp;
  Leaf<A> lr;
  lp.load(a.force, a.vel_pos, i, m, s.dt);
  lr.load(a.torque, a.vel_rot, i, m, s.dt);
  float up[D], ur[A];
  lp.mix(s, a.vel_pos, i, up);
  lr.mix(s, a.vel_rot, i, ur);
  // updates * mask, then where(mask, params + updates, params)
  if (!fixed)
#pragma unroll
    for (int c = 0; c < D; ++c) a.pos[D * i + c] = a.pos[D * i + c] + up[c] * m;
  float rv[A];
#pragma unroll
  for (int c = 0; c < A; ++c) rv[c] = fixed ? 0.0f : 0.0f + ur[c] * m;
  const Vec<3> r3 = (D == 3) ? Vec<3>{rv[0], rv[A > 1 ? 1 : 0], rv[A > 2 ? 2 : 0]} : Vec<3>{0.0f, 0.0f, rv[0]};
  const float4 q4 = reinterpret_cast<const float4*>(a.q)[i];
  const Quat q = {q4.x, q4.y, q4.z, q4.w};
  const Quat qn = q_unit(q_mul(q_from_rotvec(r3), q));
  reinterpret_cast<float4*>(a.q)[i] = make_float4(qn.w, qn.x, qn.y, qn.z);
}

__global__ void k_fire_commit(FireArgs a) {
  if (threadIdx.x != 0 || (a.stop && *a.stop)) return;
  a.st->dt = a.st->new_dt;
  a.st->alpha = a.st->new_alpha;
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_fire_grad_max(FireArgs a) {
  if (a.stop && *a.stop) return;
  constexpr int A = D == 3 ? 3 : 1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float m = 0.0f;
  if (i < a.n) {
#pragma unroll
    for (int c = 0; c < D; ++c) m = fmaxf(m, fabsf(a.force[D * i + c]));
#pragma unroll
    for (int c = 0; c < A; ++c) m = fmaxf(m, fabsf(a.torque[A * i + c]));
    // a NaN component makes the maximum NaN, as jnp.max does
    for (int c = 0; c < D; ++c)
      if (a.force[D * i + c] != a.force[D * i + c]) m = __int_as_float(0x7fc00000);
    for (int c = 0; c < A; ++c)
      if (a.torque[A * i + c] != a.torque[A * i + c]) m = __int_as_float(0x7fc00000);
  }
  // non-negative floats (and the quiet NaN above) order like their bits;
  // one atomic per block
  __shared__ unsigned wmax[kBlock / 64];
  unsigned bits = (unsigned)__float_as_int(m);
  for (int off = 32; off > 0; off >>= 1) bits = max(bits, (unsigned)__shfl_xor((int)bits, off));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned b = wmax[0];
    for (int w = 1; w < kBlock / 64; ++w) b = max(b, wmax[w]);
    atomicMax(&a.st->max_grad_bits, b);
  }
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_fire_gravity_energy(FireArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float gx;
  if (D == 3)
    gx = (a.g[0] * a.pos[3 * i] + a.g[1] * a.pos[3 * i + 1]) + a.g[2] * a.pos[3 * i + 2];
  else
    gx = a.g[0] * a.pos[2 * i] + a.g[1] * a.pos[2 * i + 1];
  a.e_out[i] += (double)(-(gx * a.mass[i]));
}

// ---- precision form ---------------------------------------------------------
// Positions are x = c h + o per axis, c an integer cell and o in [0, h): the
// float32 offset keeps ~1e-7 h of resolution anywhere in the box, where an
// absolute float32 coordinate keeps ~1e-7 L.  The cell edge is the float pair
// h + hlo, so the lattice matches the box to float64 accuracy.  The
// displacement of two spheres in neighbouring cells, (c_i - c_j) h + (o_i -
// o_j), is formed as a float pair too (error-free sums), so it is exact to
// ~1e-14 |r| rather than rounded to ~6e-8 |r|.

// x = c h + o with o folded back into [0, h) (c wrapped periodically);
// returns whether the cell changed
template <int D>
__device__ __forceinline__ bool prec_fold(const PrecArgs& a, int64_t i) {
  bool moved = false;
#pragma unroll
  for (int q = 0; q < D; ++q) {
    float o = a.off[D * i + q];
    if (!(o >= 0.0f && o < a.h[q])) {
      int c = a.cidx[D * i + q];
      const float m = floorf(o / a.h[q]);
      o = (o - m * a.h[q]) - m * a.hlo[q];
      if (o < 0.0f) o = 0.0f;  // rounding of o - m h
      if (o >= a.h[q]) o = 0.0f, c += 1;
      c = (c + (int)m) % a.dims[q];
      if (c < 0) c += a.dims[q];
      a.cidx[D * i + q] = c;
      a.off[D * i + q] = o;
      moved = true;
    }
  }
  return moved;
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_prec_keys(PrecArgs a) {
  if ((a.gated && *a.moved == 0) || (a.stop && *a.stop)) return;  // no sphere changed cell: the last sort stands
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  if (a.vl) prec_fold<D>(a, i);  // list mode: cells are kept between list builds
  const int32_t* c = a.cidx + D * i;
  a.keys[i] = D == 3 ? (uint32_t)((c[0] * a.dims[1] + c[1]) * a.dims[2] + c[2]) : (uint32_t)(c[0] * a.dims[1] + c[1]);
}

// (offset, radius) of every sphere in cell order, so the stencil reads each
// cell's candidates as one contiguous run
template <int D>
__global__ void __launch_bounds__(kBlock) k_prec_gather(PrecArgs a) {
  if (a.stop && *a.stop) return;
  const int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (kk == 0) {  // after this evaluation's sort, before its max|F| reduction
    if (a.gmax_bits) *a.gmax_bits = 0u;
    if (a.gated && !a.vl) *a.moved = 0;
  }
  if (kk >= a.n) return;
  const int64_t j = a.perm[kk];
  a.sp[kk] = make_float4(a.off[D * j], a.off[D * j + 1], D == 3 ? a.off[D * j + 2] : 0.0f, a.rad[j]);
}

// Verlet list of the precision form (rebuilds only, gated by *moved): the
// partners of the sphere in sorted slot kk0 within r_i + r_j + skin, in
// stencil order (half-width vsr <= 2), each entry the partner's sorted slot
// (low 24 bits) and the cell difference c_i - c_j (high 8 bits, base-5
// digits + 2); the build offsets.  Cells stay fixed until the next build
// (offsets may leave [0, h) by up to skin / 2 meanwhile), so the entry's
// cell difference holds for the list's lifetime.
template <int D>
__global__ void __launch_bounds__(kBlock) k_prec_vbuild(PrecArgs a) {
  if (*a.moved == 0 || (a.stop && *a.stop)) return;
  const int64_t kk0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (kk0 >= a.n) return;
  const int64_t i = a.perm[kk0];
  const float4 si = a.sp[kk0];
  const float oi[3] = {si.x, si.y, si.z};
  int ci[3] = {0, 0, 0};
  for (int q = 0; q < D; ++q) {
    ci[q] = a.cidx[D * i + q];
    a.boff[D * i + q] = oi[q];
  }
  int32_t* row = a.vl + kk0 * a.vm;
  int cnt = 0;
  const int sr = a.vsr, zr = D == 3 ? sr : 0;
  for (int ox = -sr; ox <= sr; ++ox)
    for (int oy = -sr; oy <= sr; ++oy)
      for (int oz = -zr; oz <= zr; ++oz) {
        const int o[3] = {ox, oy, oz};
        int cc[3] = {0, 0, 0};
        for (int q = 0; q < D; ++q) {
          cc[q] = (ci[q] + o[q]) % a.dims[q];
          if (cc[q] < 0) cc[q] += a.dims[q];
        }
        const int64_t h = D == 3 ? ((int64_t)cc[0] * a.dims[1] + cc[1]) * a.dims[2] + cc[2]
                                 : (int64_t)cc[0] * a.dims[1] + cc[1];
        for (int kk = a.starts[h]; kk < a.starts[h + 1]; ++kk) {
          if (kk == kk0) continue;
          const float4 sj = a.sp[kk];
          const float osj[3] = {sj.x, sj.y, sj.z};
          float d2 = 0.0f;
          for (int q = 0; q < D; ++q) {
            const float r = (float)(-o[q]) * a.h[q] + (oi[q] - osj[q]);
            d2 = d2 + r * r;
          }
          const float R = (si.w + sj.w) + a.vcut;
          if (!(d2 < R * R)) continue;
          const int code = (2 - ox) + 5 * ((2 - oy) + 5 * (2 - oz));
          if (cnt < a.vm) row[cnt] = (int32_t)((uint32_t)kk | ((uint32_t)code << 24));
          ++cnt;
        }
      }
  for (int t = cnt; t < a.vm; ++t) row[t] = -1;
  if (cnt > a.vm) *a.vovf = 1;
}

// spring law (spring.py:97-108, 138-147) on the cell stencil; float32 pair
// arithmetic, float64 energy and virial.  Thread kk takes the kk-th sphere
// in cell order, so a wavefront's spheres are spatial neighbours and their
// candidates (the cell-sorted (offset, radius) copy) are shared in cache;
// each sphere sums its candidates in stencil order, as a thread per sphere
// in any order would.
template <int D>
__global__ void __launch_bounds__(kBlock) k_prec_force(PrecArgs a) {
  if (a.stop && *a.stop) return;
  const int64_t kk0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (kk0 >= a.n) return;
  const int64_t i = a.perm[kk0];
  int ci[3] = {0, 0, 0};
  float oi[3] = {0, 0, 0};
  const float4 si = a.sp[kk0];
  const float osi[3] = {si.x, si.y, si.z};
  for (int q = 0; q < D; ++q) {
    ci[q] = a.cidx[D * i + q];
    oi[q] = osi[q];
  }
  const float ri = si.w;
  float F[3] = {0, 0, 0};
  double e = 0.0, w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int zr = D == 3 ? 1 : 0;
  for (int ox = -1; ox <= 1; ++ox)
    for (int oy = -1; oy <= 1; ++oy)
      for (int oz = -zr; oz <= zr; ++oz) {
        const int o[3] = {ox, oy, oz};
        int cc[3] = {0, 0, 0};
        for (int q = 0; q < D; ++q) {
          cc[q] = (ci[q] + o[q]) % a.dims[q];
          if (cc[q] < 0) cc[q] += a.dims[q];
        }
        const int64_t h = D == 3 ? ((int64_t)cc[0] * a.dims[1] + cc[1]) * a.dims[2] + cc[2]
                                 : (int64_t)cc[0] * a.dims[1] + cc[1];
        const int k1 = a.starts[h + 1];
        for (int kk = a.starts[h]; kk < k1; ++kk) {
          if (kk == kk0) continue;
          const float4 sj = a.sp[kk];
          const float osj[3] = {sj.x, sj.y, sj.z};
          // r = rh + rl, both float32: two_sum of the cell term (exact, the
          // stencil offset is -1, 0 or 1) and the offset difference, whose
          // own rounding error two_sum recovers too.  The contact test reads
          // rh only, so the low parts are formed for contacts alone.
          float r[3] = {0, 0, 0}, rl[3] = {0, 0, 0};
          for (int q = 0; q < D; ++q) r[q] = (float)(-o[q]) * a.h[q] + (oi[q] - osj[q]);
          const float d2 = D == 3 ? (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2] : r[0] * r[0] + r[1] * r[1];
          const float R = ri + sj.w;
          if (!(d2 < R * R) || d2 == 0.0f) continue;
          for (int q = 0; q < D; ++q) {
            const float oj = osj[q];
            const float b = oi[q] - oj;
            const float bv = b - oi[q];
            const float be = (oi[q] - (b - bv)) + (-oj - bv);
            const float c = (float)(-o[q]) * a.h[q];
            const float t = c + b;
            const float tv = t - c;
            const float te = (c - (t - tv)) + (b - tv);
            rl[q] = (te + be) + (float)(-o[q]) * a.hlo[q];
          }
          const float inv = 1.0f / sqrtf(d2);
          // delta = (R^2 - |r|^2) / (R + |r|) with |r|^2 = sum rh^2 + 2 rh rl
          // and the squares split exactly into hi + lo with fma, so the
          // cancellation of R against |r| at small overlaps loses nothing;
          // R - |r| directly would keep only ~1e-7 |r| / delta of relative
          // accuracy
          float sh = 0.0f, sl = 0.0f;
          for (int q = 0; q < D; ++q) {
            const float p = r[q] * r[q];
            const float pe = fmaf(r[q], r[q], -p);
            const float t = sh + p;
            const float bp = t - sh;
            sl = sl + ((((sh - (t - bp)) + (p - bp)) + pe) + 2.0f * r[q] * rl[q]);
            sh = t;
          }
          const float RR = R * R;
          const float diff = (RR - sh) + (fmaf(R, R, -RR) - sl);
          const float delta = diff / (R + d2 * inv);
          const float s = (a.k * delta) * inv;
          for (int q = 0; q < D; ++q) F[q] = F[q] + s * r[q];
          const double dd = (double)delta;
          e += 0.25 * (double)a.k * dd * dd;
          if (a.w_out)
            for (int p = 0; p < D; ++p)
              for (int q = 0; q < D; ++q) w[D * p + q] += 0.5 * (double)r[p] * (double)(s * r[q]);
        }
      }
  for (int q = 0; q < D; ++q) a.force[D * i + q] = F[q];
  a.e_out[i] = e;
  if (a.w_out)
    for (int q = 0; q < D * D; ++q) a.w_out[(int64_t)D * D * i + q] = w[q];
}

// The spring force of k_prec_force over the Verlet list, the pair's cell
// difference read from the entry.  With h >= 2 r_max and offsets within
// skin / 2 < h / 2 of [0, h), a contact has |c_i - c_j| <= 2 per axis, so
// (c_i - c_j) h is still exact in float32.  Partners are summed in the
// list's order (the stencil order of the build).
template <int D>
__global__ void __launch_bounds__(kBlock) k_prec_vforce(PrecArgs a) {
  if (a.stop && *a.stop) return;
  const int64_t kk0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (kk0 == 0) *a.moved = 0;  // the list of this evaluation is built
  if (kk0 >= a.n) return;
  const int64_t i = a.perm[kk0];
  const float4 si = a.sp[kk0];
  const float oi[3] = {si.x, si.y, si.z};
  const float ri = si.w;
  float F[3] = {0, 0, 0};
  double e = 0.0, w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int32_t* row = a.vl + kk0 * a.vm;
  for (int t = 0; t < a.vm; ++t) {
    const int32_t ent = row[t];
    if (ent < 0) break;
    const float4 sj = a.sp[ent & 0xFFFFFF];
    const float osj[3] = {sj.x, sj.y, sj.z};
    const int code = (int)((uint32_t)ent >> 24);
    const int dc[3] = {code % 5 - 2, (code / 5) % 5 - 2, code / 25 - 2};
    float r[3] = {0, 0, 0}, rl[3] = {0, 0, 0};
    for (int q = 0; q < D; ++q) r[q] = (float)dc[q] * a.h[q] + (oi[q] - osj[q]);
    const float d2 = D == 3 ? (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2] : r[0] * r[0] + r[1] * r[1];
    const float R = ri + sj.w;
    if (!(d2 < R * R) || d2 == 0.0f) continue;
    for (int q = 0; q < D; ++q) {
      const float oj = osj[q];
      const float b = oi[q] - oj;
      const float bv = b - oi[q];
      const float be = (oi[q] - (b - bv)) + (-oj - bv);
      const float c = (float)dc[q] * a.h[q];
      const float tt = c + b;
      const float tv = tt - c;
      const float te = (c - (tt - tv)) + (b - tv);
      rl[q] = (te + be) + (float)dc[q] * a.hlo[q];
    }
    const float inv = 1.0f / sqrtf(d2);
    float sh = 0.0f, sl = 0.0f;
    for (int q = 0; q < D; ++q) {
      const float p = r[q] * r[q];
      const float pe = fmaf(r[q], r[q], -p);
      const float tt = sh + p;
      const float bp = tt - sh;
      sl = sl + ((((sh - (tt - bp)) + (p - bp)) + pe) + 2.0f * r[q] * rl[q]);
      sh = tt;
    }
    const float RR = R * R;
    const float diff = (RR - sh) + (fmaf(R, R, -RR) - sl);
    const float delta = diff / (R + d2 * inv);
    const float s = (a.k * delta) * inv;
    for (int q = 0; q < D; ++q) F[q] = F[q] + s * r[q];
    const double dd = (double)delta;
    e += 0.25 * (double)a.k * dd * dd;
    if (a.w_out)
      for (int p = 0; p < D; ++p)
        for (int q = 0; q < D; ++q) w[D * p + q] += 0.5 * (double)r[p] * (double)(s * r[q]);
  }
  for (int q = 0; q < D; ++q) a.force[D * i + q] = F[q];
  a.e_out[i] = e;
  if (a.w_out)
    for (int q = 0; q < D * D; ++q) a.w_out[(int64_t)D * D * i + q] = w[q];
}

// the FIRE update of the position leaf (no rotation: the spring law has no
// torque), then the offset is folded back into [0, h)
template <int D>
__global__ void __launch_bounds__(kBlock) k_prec_move(FireArgs f, PrecArgs a) {
  if (a.stop && *a.stop) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const FireDev s = *f.st;
  Leaf<D> lp;
  lp.load(f.force, f.vel_pos, i, 1.0f, s.dt);
  float up[D];
  lp.mix(s, f.vel_pos, i, up);
  if (a.vl) {  // list mode: cells stay; a rebuild once the sphere has moved skin / 2
    float d2 = 0.0f;
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const float o = a.off[D * i + q] + up[q];
      a.off[D * i + q] = o;
      const float d = o - a.boff[D * i + q];
      d2 = d2 + d * d;
    }
    if (d2 > a.vtrig) *a.moved = 1;
    return;
  }
  bool moved = false;
#pragma unroll
  for (int q = 0; q < D; ++q) {
    float o = a.off[D * i + q] + up[q];
    int c = a.cidx[D * i + q];
    if (!(o >= 0.0f && o < a.h[q])) {
      const float m = floorf(o / a.h[q]);
      o = (o - m * a.h[q]) - m * a.hlo[q];
      if (o < 0.0f) o = 0.0f;  // rounding of o - m h
      if (o >= a.h[q]) o = 0.0f, c += 1;
      c = (c + (int)m) % a.dims[q];
      if (c < 0) c += a.dims[q];
      a.cidx[D * i + q] = c;
      moved = true;
    }
    a.off[D * i + q] = o;
  }
  if (moved) *a.moved = 1;
}

// The tests at the top of one iteration of minimize's loop
// (routines.py:284-314: pe_tol on the energy per sphere, pe_diff_tol on the
// relative change, force_tol on max|F|, max_steps) plus a stall test (no new
// low of max|F|, or of the energy by more than 1e-13 of it, for stall_steps
// iterations): sets stop, or counts the update about to run.
__global__ void k_prec_check(PrecStop* ps, const double* pe_sum, const unsigned* gmax_bits, PrecStopP p) {
  if (threadIdx.x != 0 || ps->stop) return;
  const double pe = *pe_sum;
  float gf;
  const unsigned bits = *gmax_bits;
  gf = __uint_as_float(bits);
  const double g = (double)gf;
  ps->pe = pe;
  ps->gmax = g;
  if (ps->first) {
    ps->best_g = g;
    ps->best_e = pe;
    ps->best_at = 0;
    ps->first = 0;
  }
  if (ps->steps >= p.max_steps) {
    ps->stop = 1;
    return;
  }
  const double tiny = 2.2250738585072014e-308;
  const bool conv_pe = fabs(pe / p.dN) <= p.pe_tol;
  const double denom = fmax(fmax(fabs(pe), fabs(ps->prev)), tiny);
  const bool conv_rel = fabs(pe - ps->prev) / denom < p.pe_diff_tol;
  if (g < ps->best_g) ps->best_g = g, ps->best_at = ps->steps;
  if (pe < ps->best_e - 1e-13 * fabs(ps->best_e)) ps->best_e = pe, ps->best_at = ps->steps;
  const bool stalled = p.stall_steps > 0 && ps->steps - ps->best_at >= p.stall_steps;
  if (conv_pe || conv_rel || g <= p.force_tol || stalled) {
    ps->stop = 1;
    return;
  }
  ps->prev = pe;
  ps->steps += 1;
}

}  // namespace

hipError_t launch_prec_check(PrecStop* ps, const double* pe_sum, const unsigned* gmax_bits, const PrecStopP& p,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_prec_check, dim3(1), dim3(64), 0, s, ps, pe_sum, gmax_bits, p);
  return hipGetLastError();
}

hipError_t launch_prec_keys(const PrecArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.dim == 3)
    hipLaunchKernelGGL(k_prec_keys<3>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_prec_keys<2>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_prec_gather(const PrecArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.dim == 3)
    hipLaunchKernelGGL(k_prec_gather<3>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_prec_gather<2>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_prec_force(const PrecArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.dim == 3)
    hipLaunchKernelGGL(k_prec_force<3>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_prec_force<2>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_prec_vbuild(const PrecArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.dim == 3)
    hipLaunchKernelGGL(k_prec_vbuild<3>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_prec_vbuild<2>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_prec_vforce(const PrecArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.dim == 3)
    hipLaunchKernelGGL(k_prec_vforce<3>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_prec_vforce<2>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_prec_move(const FireArgs& f, const PrecArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.dim == 3)
    hipLaunchKernelGGL(k_prec_move<3>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, f, a);
  else
    hipLaunchKernelGGL(k_prec_move<2>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, f, a);
  return hipGetLastError();
}

hipError_t launch_fire_power(const FireArgs& a, hipStream_t s) {
  if (a.dim == 3)
    hipLaunchKernelGGL(k_fire_power<3>, dim3(a.nparts), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_fire_power<2>, dim3(a.nparts), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_fire_decide(const FireArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_fire_decide, dim3(1), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_fire_move(const FireArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.dim == 3)
    hipLaunchKernelGGL(k_fire_move<3>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_fire_move<2>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_fire_commit(const FireArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_fire_commit, dim3(1), dim3(64), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_fire_grad_max(const FireArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.dim == 3)
    hipLaunchKernelGGL(k_fire_grad_max<3>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_fire_grad_max<2>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_fire_gravity_energy(const FireArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.dim == 3)
    hipLaunchKernelGGL(k_fire_gravity_energy<3>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_fire_gravity_energy<2>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace jdem
