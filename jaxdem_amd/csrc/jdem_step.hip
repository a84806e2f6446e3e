// jdem_step.hip -- (b) neighbour traversal with the pair laws and (c) the
// fused integration of one DEM time step on gfx950.
//
// One thread owns one particle i and accumulates F_i and T_i in registers
// over its candidates in the reference's traversal order -- no atomics, so
// results are deterministic run to run.  The same kernel then applies
// ForceManager, step_after_force of step k, and domain.apply +
// step_before_force of step k+1 (system.py:60-82), writing positions and
// velocities to the other half of a ping-pong pair because neighbours of
// i still read the step-k values.
//
// Reference: colliders/naive.py:190-235, cell_list.py:188-261, 437-496,
// neighbor_list.py:594-669; forces/force_manager.py:341-425;
// integrators/velocity_verlet.py:57-94, direct_euler.py:62-65,
// velocity_verlet_spiral.py:83-180, spiral.py:60-144; domains/reflect.py:
// 102-297.
#include <hip/hip_runtime.h>

#include "jdem_internal.h"

namespace jdem {

template <int D>
JD Vec<D> v4s(float4 a);
template <>
JD Vec<3> v4s<3>(float4 a) { return {a.x, a.y, a.z}; }
template <>
JD Vec<2> v4s<2>(float4 a) { return {a.x, a.y}; }

template <class Real>
JD bool is_bonded(const StepArgsT<Real>& a, int64_t i, int32_t j) {
  for (int w = 0; w < a.bond_w; ++w)
    if (a.bond_id[i * a.bond_w + w] == j) return true;
  return false;
}

template <int LAW, class Real>
JD int species_of(const LawPT<Real>& L, int64_t i) {
  if constexpr (LawTraits<LAW>::routed) return L.species ? L.species[i] : 0;
  return 0;
}

template <int D, int LAW, class Real>
JD void fill_partner(const StepArgsT<Real>& a, Side<D, Real>& b, int32_t j) {
  if (a.law.n_mat > 1) b.mat = a.mat_id[j];
  b.sp = species_of<LAW>(a.law, j);
  if constexpr (LawTraits<LAW>::kin) {
    b.v = ld<D>(a.vel_in, j);
    b.w = lda<D>(a.ang_in, j);
    b.m = a.mass[j];
  }
}

// Evaluate the pair (i <- j) and accumulate.  Returns true if in contact.
template <int D, int LAW, bool PER, class Real>
JD bool visit(const StepArgsT<Real>& a, const Side<D, Real>& me, int64_t i, int32_t j, const Vec<D, Real>& xj,
              Real rj, Vec<D, Real>& F, Ang<D, Real>& T, const Vec<D, Real>& xi_old, Vec<D, Real>& xi_new) {
  const Vec<D, Real> rij = disp_<PER>(me.x, xj, a.dom);
  const Real d_sq = dot(rij, rij);
  const Real R = me.r + rj;
  // No contact is possible: delta of every law is exactly 0 (|rij| exceeds
  // R by far more than the rounding of d_sq * rsqrt(d_sq)).
  if (d_sq > (R * R) * Real(1.0001)) return false;
  if (a.clump ? a.clump[j] == a.clump[i] : j == (int32_t)i) return false;  // valid_interaction_mask
  if (!a.isb && a.bond_w > 0 && is_bonded(a, i, j)) return false;
  Side<D, Real> b;
  b.x = xj;
  b.r = rj;
  b.mat = 0;
  fill_partner<D, LAW>(a, b, j);
  PairOut<D, Real> o = law_eval<D, LAW>(a.law, me, b, rij, d_sq, xi_old, a.dt);
  F = F + o.F;
  if constexpr (LawTraits<LAW>::torque) T = T + o.T;
  if constexpr (LawTraits<LAW>::hist) xi_new = o.xi;
  return o.contact;
}

// Per-particle force from the three candidate sources.
template <int D, int LAW, int SRC, bool PER, class Real>
JD void particle_force(const StepArgsT<Real>& a, int64_t idx, int64_t i, const Side<D, Real>& me, Vec<D, Real>& F,
                       Ang<D, Real>& T) {
  F = zero_vec<D, Real>();
  T = zero_ang<D, Real>();
  Vec<D, Real> xo = zero_vec<D, Real>(), xn;
  if constexpr (SRC == SRC_NAIVE) {
    for (int64_t j = 0; j < a.n; ++j) {
      visit<D, LAW, PER>(a, me, i, (int32_t)j, ld<D>(a.pos_in, j), a.rad[j], F, T, xo, xn);
    }
  } else if constexpr (SRC == SRC_CELLS) {
    const uint32_t key = (uint32_t)a.skeys[idx];
    int c[3];
    c[0] = (int)(key % (uint32_t)a.grid.dims[0]);
    c[1] = (int)((key / (uint32_t)a.grid.dims[0]) % (uint32_t)a.grid.dims[1]);
    c[2] = (D == 3) ? (int)(key / ((uint32_t)a.grid.dims[0] * (uint32_t)a.grid.dims[1])) : 0;
    const int sr = a.grid.sr;
    int hi[3];
    for (int ax = 0; ax < D; ++ax) hi[ax] = PER ? min(sr, a.grid.dims[ax] - sr - 1) : sr;
    for (int ox = -sr; ox <= hi[0]; ++ox) {
      int cx = c[0] + ox;
      if (PER) {
        cx = cx % a.grid.dims[0];
        if (cx < 0) cx += a.grid.dims[0];
      } else if (cx < 0 || cx >= a.grid.dims[0])
        continue;
      for (int oy = -sr; oy <= hi[1]; ++oy) {
        int cy = c[1] + oy;
        if (PER) {
          cy = cy % a.grid.dims[1];
          if (cy < 0) cy += a.grid.dims[1];
        } else if (cy < 0 || cy >= a.grid.dims[1])
          continue;
        const int zlo = (D == 3) ? -sr : 0, zhi = (D == 3) ? hi[2] : 0;
        for (int oz = zlo; oz <= zhi; ++oz) {
          int cz = 0;
          if (D == 3) {
            cz = c[2] + oz;
            if (PER) {
              cz = cz % a.grid.dims[2];
              if (cz < 0) cz += a.grid.dims[2];
            } else if (cz < 0 || cz >= a.grid.dims[2])
              continue;
          }
          const int h = cx * a.grid.strides[0] + cy * a.grid.strides[1] + (D == 3 ? cz * a.grid.strides[2] : 0);
          const int k1 = a.cend[h];
          for (int k = a.cstart[h]; k < k1; ++k) {
            if constexpr (std::is_same<Real, float>::value) {
              const float4 p = a.spos[k];
              visit<D, LAW, PER>(a, me, i, a.perm[k], v4s<D>(p), p.w, F, T, xo, xn);
            } else {  // float64: the partner's own arrays
              const int32_t j = a.perm[k];
              visit<D, LAW, PER>(a, me, i, j, ld<D>(a.pos_in, j), a.rad[j], F, T, xo, xn);
            }
          }
        }
      }
    }
  } else {  // SRC_NLIST
    const int32_t* row = a.nl + i * a.M;
    for (int w = 0; w < a.W32; ++w) {
      uint32_t live_old = 0, live_new = 0;
      if constexpr (LawTraits<LAW>::hist) live_old = a.hist_live[i * a.W32 + w];
      const int s_end = min(a.M, (w + 1) * 32);
      bool done = false;
      for (int sl = w * 32; sl < s_end; ++sl) {
        const int32_t j = row[sl];
        if (j < 0) {
          done = true;
          break;
        }
        Vec<D, Real> xold = zero_vec<D, Real>();
        if constexpr (LawTraits<LAW>::hist) {
          if ((live_old >> (sl & 31)) & 1u) xold = ld<D>(a.hist_xi, i * a.M + sl);
        }
        Vec<D, Real> xnew;
        const bool c = visit<D, LAW, PER>(a, me, i, j, ld<D>(a.pos_in, j), a.rad[j], F, T, xold, xnew);
        if constexpr (LawTraits<LAW>::hist) {
          if (c) {
            st(a.hist_xi, i * a.M + sl, xnew);
            live_new |= 1u << (sl & 31);
          }
        }
      }
      if constexpr (LawTraits<LAW>::hist) a.hist_live[i * a.W32 + w] = live_new;
      if (done) {
        if constexpr (LawTraits<LAW>::hist)
          for (int w2 = w + 1; w2 < a.W32; ++w2) a.hist_live[i * a.W32 + w2] = 0u;
        break;
      }
    }
  }
}

template <int D, int LAW, class Real>
JD void load_side(const StepArgsT<Real>& a, int64_t i, Side<D, Real>& me) {
  me.x = ld<D>(a.pos_in, i);
  me.r = a.rad[i];
  me.mat = (a.law.n_mat > 1) ? a.mat_id[i] : 0;
  me.sp = species_of<LAW>(a.law, i);
  me.v = ld<D>(a.vel_in, i);
  me.w = lda<D>(a.ang_in, i);
  me.m = a.mass[i];
}

// The fused step kernel.  Phases (flags): do_force computes F, T at the
// current positions (else they are read from force/torque); do_after runs
// ForceManager + step_after_force of step k; do_before runs domain.apply +
// step_before_force of step k+1 and the NeighborList displacement check.
template <int D, int LAW, int SRC, bool PER, class Real>
__global__ void __launch_bounds__(kBlock) k_step(StepArgsT<Real> a) {
  if (a.flag_rd && *a.flag_rd) {  // a neighbour-list rebuild is pending
    if (a.flag_wr) *a.flag_wr = *a.flag_rd;
    return;
  }
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.n) return;
  const int64_t i = (SRC == SRC_CELLS && a.do_force) ? (int64_t)a.perm[idx] : idx;
  Side<D, Real> me;
  load_side<D, LAW>(a, i, me);
  Vec<D, Real> F;
  Ang<D, Real> T;
  if (a.do_force && a.force_zero) {
    F = zero_vec<D, Real>();
    T = zero_ang<D, Real>();
  } else if (a.do_force) {
    particle_force<D, LAW, SRC, PER>(a, idx, i, me, F, T);
    // clump member: the lever torque of its summed force (naive.py:231,
    // cell_list.py:462)
    if (a.ppr) T = T + cross_t(ld<D>(a.ppr, i), F);
  } else {
    F = ld<D>(a.force, i);
    T = lda<D>(a.torque, i);
  }
  if (!a.write_kin) {  // Collider.compute_force only
    st(a.force, i, F);
    sta(a.torque, i, T);
    return;
  }
  Vec<D, Real> x = me.x, v = me.v;
  Ang<D, Real> w = me.w;
  QuatT<Real> q = ldq(a.q, i);
  const Ang<D, Real> I = lda<D>(a.inertia, i);
  const Real m = me.m;
  const bool fixed = a.fixed[i] != 0;
  const Real freef = fixed ? Real(0) : Real(1);
  if (a.do_after) {
    // ForceManager.apply for spheres (force_manager.py:358-424)
    if (a.has_fm) {
      Vec<D, Real> fp = a.ext_f ? ld<D>(a.ext_f, i) : zero_vec<D, Real>();
      Vec<D, Real> fc = a.ext_fc ? ld<D>(a.ext_fc, i) : zero_vec<D, Real>();
      Vec<D, Real> g;
      if constexpr (D == 3)
        g = Vec<3, Real>{a.g[0], a.g[1], a.g[2]};
      else
        g = Vec<2, Real>{a.g[0], a.g[1]};
      fc = fc + g * (m / Real(1));
      F = (F + fp) + fc;
      if (a.ext_t) T = T + lda<D>(a.ext_t, i);
    }
    // step_after_force
    if (a.lin == JDEM_LIN_VERLET) {
      const Real s = (a.dt * Real(0.5)) / m;
      v = v + (F * s) * freef;
    } else if (a.lin == JDEM_LIN_EULER) {
      const Real s = a.dt / m;
      v = v + (F * s) * freef;
      x = x + v * a.dt;
    }
    if (a.rot == JDEM_ROT_VERLET_SPIRAL)
      vvs_after<D>(q, w, T, I, a.dt, freef);
    else if (a.rot == JDEM_ROT_SPIRAL)
      spiral_after<D>(q, w, T, I, a.dt, freef);
  }
  if (a.do_before) {
    // domain.apply of step k+1 (reflect.py:102-297; free/periodic: no-op)
    if (a.dom.kind == JDEM_DOMAIN_REFLECT) reflect_apply<D>(a.dom, x, v, q, F, m, me.r, fixed, a.dt);
    // step_before_force of step k+1
    if (a.lin == JDEM_LIN_VERLET) {
      const Real s = (a.dt * Real(0.5)) / m;
      v = v + (F * s) * freef;
      x = x + v * a.dt;
    }
    if (a.rot == JDEM_ROT_VERLET_SPIRAL) vvs_before<D>(q, w, T, I, a.dt, freef);
  }
  st(a.pos_out, i, x);
  st(a.vel_out, i, v);
  sta(a.ang_out, i, w);
  stq(a.q, i, q);
  st(a.force, i, F);
  sta(a.torque, i, T);
  if (a.do_before && a.check_disp) {
    const Vec<D, Real> d = x - ld<D>(a.old_pos, i);
    if (dot(d, d) > a.trigger) *a.flag_wr = a.step_idx + 1;
  }
}

// Potential energy: 1/2 sum of pair energies per particle, float64 sum.
template <int D, int LAW, int SRC, bool PER, class Real>
__global__ void __launch_bounds__(kBlock) k_energy(StepArgsT<Real> a) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.n) return;
  const int64_t i = (SRC == SRC_CELLS) ? (int64_t)a.perm[idx] : idx;
  Side<D, Real> me;
  me.x = ld<D>(a.pos_in, i);
  me.r = a.rad[i];
  me.mat = (a.law.n_mat > 1) ? a.mat_id[i] : 0;
  me.sp = species_of<LAW>(a.law, i);
  double acc = 0.0;
  auto one = [&](int32_t j, Vec<D, Real> xj, Real rj) {
    if (a.clump ? a.clump[j] == a.clump[i] : j == (int32_t)i) return;
    if (!a.isb && a.bond_w > 0 && is_bonded(a, i, j)) return;
    const Vec<D, Real> rij = disp_<PER>(me.x, xj, a.dom);
    const Real d_sq = dot(rij, rij);
    Side<D, Real> b;
    b.x = xj;
    b.r = rj;
    b.mat = (a.law.n_mat > 1) ? a.mat_id[j] : 0;
    b.sp = species_of<LAW>(a.law, j);
    const Real e = law_energy<D, LAW>(a.law, me, b, d_sq);
    acc += 0.5 * (double)e;
  };
  if constexpr (SRC == SRC_NAIVE) {
    for (int64_t j = 0; j < a.n; ++j) one((int32_t)j, ld<D>(a.pos_in, j), a.rad[j]);
  } else if constexpr (SRC == SRC_NLIST) {
    const int32_t* row = a.nl + i * a.M;
    for (int sl = 0; sl < a.M; ++sl) {
      const int32_t j = row[sl];
      if (j < 0) break;
      one(j, ld<D>(a.pos_in, j), a.rad[j]);
    }
  } else {
    const uint32_t key = (uint32_t)a.skeys[idx];
    int c[3];
    c[0] = (int)(key % (uint32_t)a.grid.dims[0]);
    c[1] = (int)((key / (uint32_t)a.grid.dims[0]) % (uint32_t)a.grid.dims[1]);
    c[2] = (D == 3) ? (int)(key / ((uint32_t)a.grid.dims[0] * (uint32_t)a.grid.dims[1])) : 0;
    const int sr = a.grid.sr;
    int hi[3];
    for (int ax = 0; ax < D; ++ax) hi[ax] = PER ? min(sr, a.grid.dims[ax] - sr - 1) : sr;
    int o[3];
    for (o[0] = -sr; o[0] <= hi[0]; ++o[0])
      for (o[1] = -sr; o[1] <= hi[1]; ++o[1])
        for (o[2] = (D == 3 ? -sr : 0); o[2] <= (D == 3 ? hi[2] : 0); ++o[2]) {
          int h = 0;
          bool ok = true;
          for (int ax = 0; ax < D; ++ax) {
            int cc = c[ax] + o[ax];
            if (PER) {
              cc = cc % a.grid.dims[ax];
              if (cc < 0) cc += a.grid.dims[ax];
            } else if (cc < 0 || cc >= a.grid.dims[ax]) {
              ok = false;
            }
            h += cc * a.grid.strides[ax];
          }
          if (!ok) continue;
          for (int k = a.cstart[h]; k < a.cend[h]; ++k) {
            if constexpr (std::is_same<Real, float>::value) {
              const float4 p = a.spos[k];
              one(a.perm[k], v4s<D>(p), p.w);
            } else {
              const int32_t j = a.perm[k];
              one(j, ld<D>(a.pos_in, j), a.rad[j]);
            }
          }
        }
  }
  a.e_out[i] = acc;
}

// Contact virial sum_{i<j} r_ij (x) F_ij (contacts.py:89-164): each i-side
// pair contributes 1/2 r_ij (x) F_ij, float64.
template <int D, int LAW, int SRC, bool PER, class Real>
__global__ void __launch_bounds__(kBlock) k_virial(StepArgsT<Real> a) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.n) return;
  const int64_t i = (SRC == SRC_CELLS) ? (int64_t)a.perm[idx] : idx;
  Side<D, Real> me;
  load_side<D, LAW>(a, i, me);
  double W[D * D];
  for (int c = 0; c < D * D; ++c) W[c] = 0.0;
  auto one = [&](int32_t j, Vec<D, Real> xj, Real rj, Vec<D, Real> xold) {
    if (j == (int32_t)i) return;
    if (!a.isb && a.bond_w > 0 && is_bonded(a, i, j)) return;
    const Vec<D, Real> rij = disp_<PER>(me.x, xj, a.dom);
    const Real d_sq = dot(rij, rij);
    Side<D, Real> b;
    b.x = xj;
    b.r = rj;
    b.mat = 0;
    fill_partner<D, LAW>(a, b, j);
    PairOut<D, Real> o = law_eval<D, LAW>(a.law, me, b, rij, d_sq, xold, a.dt);
    const Real* r = &rij.x;
    const Real* f = &o.F.x;
    for (int p = 0; p < D; ++p)
      for (int c = 0; c < D; ++c) W[p * D + c] += 0.5 * ((double)r[p] * (double)f[c]);
  };
  if constexpr (SRC == SRC_NAIVE) {
    for (int64_t j = 0; j < a.n; ++j) one((int32_t)j, ld<D>(a.pos_in, j), a.rad[j], zero_vec<D, Real>());
  } else if constexpr (SRC == SRC_NLIST) {
    const int32_t* row = a.nl + i * a.M;
    for (int sl = 0; sl < a.M; ++sl) {
      const int32_t j = row[sl];
      if (j < 0) break;
      Vec<D, Real> xold = zero_vec<D, Real>();
      if constexpr (LawTraits<LAW>::hist) {
        if ((a.hist_live[i * a.W32 + (sl >> 5)] >> (sl & 31)) & 1u) xold = ld<D>(a.hist_xi, i * a.M + sl);
      }
      one(j, ld<D>(a.pos_in, j), a.rad[j], xold);
    }
  } else {
    const uint32_t key = (uint32_t)a.skeys[idx];
    int c[3];
    c[0] = (int)(key % (uint32_t)a.grid.dims[0]);
    c[1] = (int)((key / (uint32_t)a.grid.dims[0]) % (uint32_t)a.grid.dims[1]);
    c[2] = (D == 3) ? (int)(key / ((uint32_t)a.grid.dims[0] * (uint32_t)a.grid.dims[1])) : 0;
    const int sr = a.grid.sr;
    int hi[3];
    for (int ax = 0; ax < D; ++ax) hi[ax] = PER ? min(sr, a.grid.dims[ax] - sr - 1) : sr;
    int o[3];
    for (o[0] = -sr; o[0] <= hi[0]; ++o[0])
      for (o[1] = -sr; o[1] <= hi[1]; ++o[1])
        for (o[2] = (D == 3 ? -sr : 0); o[2] <= (D == 3 ? hi[2] : 0); ++o[2]) {
          int h = 0;
          bool ok = true;
          for (int ax = 0; ax < D; ++ax) {
            int cc = c[ax] + o[ax];
            if (PER) {
              cc = cc % a.grid.dims[ax];
              if (cc < 0) cc += a.grid.dims[ax];
            } else if (cc < 0 || cc >= a.grid.dims[ax]) {
              ok = false;
            }
            h += cc * a.grid.strides[ax];
          }
          if (!ok) continue;
          for (int k = a.cstart[h]; k < a.cend[h]; ++k) {
            if constexpr (std::is_same<Real, float>::value) {
              const float4 p = a.spos[k];
              one(a.perm[k], v4s<D>(p), p.w, zero_vec<D>());
            } else {
              const int32_t j = a.perm[k];
              one(j, ld<D>(a.pos_in, j), a.rad[j], zero_vec<D, Real>());
            }
          }
        }
  }
  for (int c = 0; c < D * D; ++c) a.e_out[i * D * D + c] = W[c];
}

// Kinetic energy per particle (utils/thermal.py:83-122), float64.
template <int D, class Real>
__global__ void __launch_bounds__(kBlock) k_kinetic(StepArgsT<Real> a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  // a clump member counts 1/(member count) of its clump (thermal.py:20-77)
  const double cnt = a.clump ? (double)(a.cl_start[a.clump[i] + 1] - a.cl_start[a.clump[i]]) : 1.0;
  const double m = (double)a.mass[i] / cnt;
  double ke = 0.0;  // rotational part
  double v2 = 0.0;
  for (int c = 0; c < D; ++c) {
    const double v = a.vel_in[D * i + c];
    v2 += v * v;
  }
  const double ke_t = 0.5 * m * v2;
  if constexpr (D == 3) {
    // body-frame angular velocity in float64
    const double qw = a.q[4 * i], qx = a.q[4 * i + 1], qy = a.q[4 * i + 2], qz = a.q[4 * i + 3];
    // rotate_back: the rotation by the conjugate (quaternion.py:267-268)
    const double cw = qw, cx = -qx, cy = -qy, cz = -qz;
    const double wx = a.ang_in[3 * i], wy = a.ang_in[3 * i + 1], wz = a.ang_in[3 * i + 2];
    const double tx = 2.0 * (cy * wz - cz * wy), ty = 2.0 * (cz * wx - cx * wz), tz = 2.0 * (cx * wy - cy * wx);
    const double bx = (wx + cw * tx) + (cy * tz - cz * ty);
    const double by = (wy + cw * ty) + (cz * tx - cx * tz);
    const double bz = (wz + cw * tz) + (cx * ty - cy * tx);
    ke += 0.5 * ((double)a.inertia[3 * i] * bx * bx + (double)a.inertia[3 * i + 1] * by * by +
                 (double)a.inertia[3 * i + 2] * bz * bz);
  } else {
    const double w = a.ang_in[i];
    ke += 0.5 * (double)a.inertia[i] * w * w;
  }
  ke = ke / cnt;
  // translational and rotational parts, summed separately
  a.e_out[2 * i] = ke_t;
  a.e_out[2 * i + 1] = ke;
}

// One stage of _step_once on its own (jdem_step_stage).
template <int D>
__global__ void __launch_bounds__(kBlock) k_stage(StepArgs a, int stage) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  Vec<D> x = ld<D>(a.pos_out, i);
  Vec<D> v = ld<D>(a.vel_out, i);
  Ang<D> w = lda<D>(a.ang_out, i);
  Quat q = ldq(a.q, i);
  Vec<D> F = ld<D>(a.force, i);
  Ang<D> T = lda<D>(a.torque, i);
  const Ang<D> I = lda<D>(a.inertia, i);
  const float m = a.mass[i];
  const bool fixed = a.fixed[i] != 0;
  const float freef = fixed ? 0.0f : 1.0f;
  switch (stage) {
    case JDEM_STAGE_DOMAIN_APPLY:
      if (a.dom.kind == JDEM_DOMAIN_REFLECT) reflect_apply<D>(a.dom, x, v, q, F, m, a.rad[i], fixed, a.dt);
      break;
    case JDEM_STAGE_LIN_BEFORE:
      if (a.lin == JDEM_LIN_VERLET) {
        const float s = (a.dt * 0.5f) / m;
        v = v + (F * s) * freef;
        x = x + v * a.dt;
      }
      break;
    case JDEM_STAGE_ROT_BEFORE:
      if (a.rot == JDEM_ROT_VERLET_SPIRAL) vvs_before<D>(q, w, T, I, a.dt, freef);
      break;
    case JDEM_STAGE_FORCE_MANAGER: {
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
      break;
    }
    case JDEM_STAGE_LIN_AFTER:
      if (a.lin == JDEM_LIN_VERLET) {
        const float s = (a.dt * 0.5f) / m;
        v = v + (F * s) * freef;
      } else if (a.lin == JDEM_LIN_EULER) {
        const float s = a.dt / m;
        v = v + (F * s) * freef;
        x = x + v * a.dt;
      }
      break;
    case JDEM_STAGE_ROT_AFTER:
      if (a.rot == JDEM_ROT_VERLET_SPIRAL)
        vvs_after<D>(q, w, T, I, a.dt, freef);
      else if (a.rot == JDEM_ROT_SPIRAL)
        spiral_after<D>(q, w, T, I, a.dt, freef);
      break;
    default:
      break;
  }
  st(a.pos_out, i, x);
  st(a.vel_out, i, v);
  sta(a.ang_out, i, w);
  stq(a.q, i, q);
  st(a.force, i, F);
  sta(a.torque, i, T);
}

// ---------------------------------------------------------------------------
// clumps (rigid bodies of spheres; every member holds its clump's pos_c, q,
// velocities, mass and inertia)
// ---------------------------------------------------------------------------
// State._pos_p_rot = R(q) pos_p and State.pos = pos_c + R(q) pos_p
// (state.py:230-304)
template <int D, class Real>
__global__ void __launch_bounds__(kBlock) k_clump_lab(StepArgsT<Real> a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const QuatT<Real> q = ldq(a.q, i);
  const Vec<D, Real> r = q_rotate(q, ld<D>(a.pos_p, i));
  st(a.ppr, i, r);
  st(a.plab, i, ld<D>(a.pos_in, i) + r);
}

// ForceManager.apply for clumps (force_manager.py:359-412): per member,
// gravity with mass / member count, the lever torque of particle forces,
// the totals; then each member's force and torque are its clump's sums of
// the totals, in ascending member index.
template <int D, class Real>
__global__ void __launch_bounds__(kBlock) k_clump_total(StepArgsT<Real> a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int c = a.clump[i];
  const Real cnt = (Real)(a.cl_start[c + 1] - a.cl_start[c]);
  const Vec<D, Real> fp = a.ext_f ? ld<D>(a.ext_f, i) : zero_vec<D, Real>();
  Vec<D, Real> fc = a.ext_fc ? ld<D>(a.ext_fc, i) : zero_vec<D, Real>();
  Ang<D, Real> tp = a.ext_t ? lda<D>(a.ext_t, i) : zero_ang<D, Real>();
  Vec<D, Real> g;
  if constexpr (D == 3)
    g = Vec<3, Real>{a.g[0], a.g[1], a.g[2]};
  else
    g = Vec<2, Real>{a.g[0], a.g[1]};
  fc = fc + g * (a.mass[i] / cnt);
  tp = tp + cross_t(ld<D>(a.ppr, i), fp);
  st(a.ftot, i, (ld<D>(a.force, i) + fp) + fc);
  sta(a.ttot, i, lda<D>(a.torque, i) + tp);
}

template <int D, class Real>
__global__ void __launch_bounds__(kBlock) k_clump_sum(StepArgsT<Real> a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int c = a.clump[i];
  Vec<D, Real> F = zero_vec<D, Real>();
  Ang<D, Real> T = zero_ang<D, Real>();
  for (int k = a.cl_start[c]; k < a.cl_start[c + 1]; ++k) {
    const int32_t m = a.cl_members[k];
    F = F + ld<D>(a.ftot, m);
    T = T + lda<D>(a.ttot, m);
  }
  st(a.force, i, F);
  sta(a.torque, i, T);
}

// ReflectDomain.apply for clumps (reflect.py:160-297).  Per member: its
// overlap past each wall at its lab position; the deepest member of its clump
// per axis takes the impulse with the effective mass of the rigid body; the
// clump's time-of-collision fraction is its members' minimum; the members'
// impulses are summed over the clump into the linear and angular velocity
// changes, and position and orientation are corrected over the remaining
// time.  Matrix products sum in index order.  Scratch per member (rscr):
// [0, D) over_lo, [D, 2D) over_hi, [2D, 3D) wall sign, [3D, 4D) active,
// [4D, 5D) denominator, [5D, 6D) contact velocity, [6D, 7D) contact
// acceleration, 7D alpha minimum, 7D + 1 clump alpha, [7D + 2, 8D + 2)
// impulse / mass, then the angular velocity change (A words).
template <int D, class Real>
JD Real* refl_row(const StepArgsT<Real>& a, int64_t i) {
  return a.rscr + (int64_t)(8 * D + 2 + (D == 3 ? 3 : 1)) * i;
}
template <class Real>
JD void to_arr_a(Vec<3, Real> v, Real* x) {
  to_arr(v, x);
}
template <class Real>
JD void to_arr_a(Ang1T<Real> v, Real* x) {
  x[0] = v.z;
}
// rows of R (body -> lab): row k = R^T e_k = rotate_back(q, e_k)
template <int D, class Real>
JD void rot_rows(QuatT<Real> q, Real R[D][D]) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    Real e[D];
#pragma unroll
    for (int c = 0; c < D; ++c) e[c] = (c == k) ? Real(1) : Real(0);
    Vec<D, Real> ek;
    from_arr(ek, e);
    Vec<D, Real> rk;
    if constexpr (D == 3)
      rk = q_rotate_back(q, ek);
    else
      rk = q_rotate(q_conj(q), ek);
    to_arr(rk, R[k]);
  }
}
// R^T v (T = true) or R v, summing over j in index order
template <int D, bool T, class Real>
JD void rmul(const Real R[D][D], const Real* v, Real* out) {
#pragma unroll
  for (int i = 0; i < D; ++i) {
    Real acc = (T ? R[0][i] : R[i][0]) * v[0];
#pragma unroll
    for (int j = 1; j < D; ++j) acc = acc + (T ? R[j][i] : R[i][j]) * v[j];
    out[i] = acc;
  }
}

template <int D, class Real>
__global__ void __launch_bounds__(kBlock) k_crefl_over(StepArgsT<Real> a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const Vec<D, Real> r = q_rotate(ldq(a.q, i), ld<D>(a.pos_p, i));
  Real pos[D];
  to_arr(ld<D>(a.pos_in, i) + r, pos);
  const Real rad = a.rad[i];
  Real* s = refl_row<D>(a, i);
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const Real lo = a.dom.anchor[c] + rad;
    const Real hi = (a.dom.anchor[c] + a.dom.box[c]) - rad;
    s[c] = fmax_(Real(0), lo - pos[c]);
    s[D + c] = fmax_(Real(0), pos[c] - hi);
  }
}

template <int D, class Real>
__global__ void __launch_bounds__(kBlock) k_crefl_alpha(StepArgsT<Real> a) {
  constexpr int A = D == 3 ? 3 : 1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int c0 = a.clump[i];
  Real mlo[D], mhi[D];
#pragma unroll
  for (int c = 0; c < D; ++c) mlo[c] = mhi[c] = -Real(INFINITY);
  for (int k = a.cl_start[c0]; k < a.cl_start[c0 + 1]; ++k) {
    const Real* sm = refl_row<D>(a, a.cl_members[k]);
#pragma unroll
    for (int c = 0; c < D; ++c) {
      mlo[c] = fmax_(mlo[c], sm[c]);
      mhi[c] = fmax_(mhi[c], sm[D + c]);
    }
  }
  Real* s = refl_row<D>(a, i);
  const QuatT<Real> q = ldq(a.q, i);
  const Vec<D, Real> pp = ld<D>(a.pos_p, i);
  const Vec<D, Real> r = q_rotate(q, pp);
  const Real inv_mass = Real(1) / a.mass[i];
  Real I[A], invI[A];
  to_arr_a(lda<D>(a.inertia, i), I);
#pragma unroll
  for (int k = 0; k < A; ++k) invI[k] = Real(1) / I[k];
  Real R[D][D];
  rot_rows<D>(q, R);
  Real delta[D], ws[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    delta[c] = fmax_(mlo[c], mhi[c]);
    const bool dlo = (s[c] > Real(0)) && (s[c] == mlo[c]);
    const bool dhi = (s[D + c] > Real(0)) && (s[D + c] == mhi[c]);
    ws[c] = (dlo ? Real(1) : Real(0)) - (dhi ? Real(1) : Real(0));
    s[2 * D + c] = ws[c];
    s[3 * D + c] = fabs_(ws[c]);
    // effective mass along axis c: 1/m + sum_k invI_k ((pos_p x n'_c)_k)^2
    Vec<D, Real> nc;
    from_arr(nc, R[c]);
    Real x[A];
    to_arr_a(cross_t(pp, nc), x);
    Real drot = (x[0] * x[0]) * invI[0];
#pragma unroll
    for (int k = 1; k < A; ++k) drot = drot + (x[k] * x[k]) * invI[k];
    Real den = inv_mass + drot;
    s[4 * D + c] = (den == Real(0)) ? Real(1) : den;
  }
  const Vec<D, Real> vcs = ld<D>(a.vel_in, i) + cross_wr(lda<D>(a.ang_in, i), r);
  const Vec<D, Real> acl = ld<D>(a.force, i) * inv_mass;
  Ang<D, Real> arot;
  if constexpr (D == 3) {
    Real t[3], w[3], tb[3], wb[3];
    to_arr(lda<3>(a.torque, i), t);
    to_arr(lda<3>(a.ang_in, i), w);
    rmul<3, true>(R, t, tb);
    rmul<3, true>(R, w, wb);
    Vec<3, Real> wbv, tbv, Iv;
    from_arr(wbv, wb);
    from_arr(tbv, tb);
    from_arr(Iv, I);
    const Vec<3, Real> ab = (tbv - cross(wbv, Iv * wbv)) * Vec<3, Real>{invI[0], invI[1], invI[2]};
    Real abv[3], ar[3];
    to_arr(ab, abv);
    rmul<3, false>(R, abv, ar);
    from_arr(arot, ar);
  } else {
    arot = lda<2>(a.torque, i) * Ang1T<Real>{invI[0]};
  }
  const Vec<D, Real> acc = acl + cross_wr(arot, r);
  Real v[D], ac[D];
  to_arr(vcs, v);
  to_arr(acc, ac);
  Real amin = Real(1);
#pragma unroll
  for (int c = 0; c < D; ++c) {
    s[5 * D + c] = v[c];
    s[6 * D + c] = ac[c];
    const Real al = toc_alpha(v[c], ac[c], delta[c], ws[c], a.dt);
    if (fabs_(ws[c]) > Real(0)) amin = fmin_(amin, al);
  }
  s[7 * D] = amin;
}

template <int D, class Real>
__global__ void __launch_bounds__(kBlock) k_crefl_impulse(StepArgsT<Real> a) {
  constexpr int A = D == 3 ? 3 : 1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int c0 = a.clump[i];
  Real acl = Real(INFINITY), cnt[D];
#pragma unroll
  for (int c = 0; c < D; ++c) cnt[c] = Real(0);
  for (int k = a.cl_start[c0]; k < a.cl_start[c0 + 1]; ++k) {
    const Real* sm = refl_row<D>(a, a.cl_members[k]);
    acl = fmin_(acl, sm[7 * D]);
#pragma unroll
    for (int c = 0; c < D; ++c) cnt[c] = cnt[c] + sm[3 * D + c];
  }
  Real* s = refl_row<D>(a, i);
  s[7 * D + 1] = acl;
  const QuatT<Real> q = ldq(a.q, i);
  const Real inv_mass = Real(1) / a.mass[i];
  Real I[A];
  to_arr_a(lda<D>(a.inertia, i), I);
  Real R[D][D];
  rot_rows<D>(q, R);
  const Real e = a.dom.restitution;
  const Real dt_factor = (acl - Real(1)) * a.dt;
  Real j[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const Real act = s[3 * D + c];
    const Real weight = act / fmax_(cnt[c], Real(1));
    const Real vc = s[5 * D + c] + dt_factor * s[6 * D + c];
    Real jj = ((-(Real(1) + e)) * vc) / s[4 * D + c];
    const Real closing = ((vc * s[2 * D + c]) < Real(0)) ? Real(1) : Real(0);
    jj = jj * (weight * closing);
    j[c] = jj;
    s[7 * D + 2 + c] = jj * inv_mass;
  }
  Real jb[D];
  rmul<D, true>(R, j, jb);
  Vec<D, Real> jbv;
  from_arr(jbv, jb);
  Real m[A], dw[A];
  to_arr_a(cross_t(ld<D>(a.pos_p, i), jbv), m);
  if constexpr (D == 3) {
    Real mb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) mb[k] = m[k] * (Real(1) / I[k]);
    rmul<3, false>(R, mb, dw);
  } else {
    dw[0] = m[0] * (Real(1) / I[0]);
  }
#pragma unroll
  for (int k = 0; k < A; ++k) s[8 * D + 2 + k] = dw[k];
}

template <int D, class Real>
__global__ void __launch_bounds__(kBlock) k_crefl_apply(StepArgsT<Real> a) {
  constexpr int A = D == 3 ? 3 : 1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int c0 = a.clump[i];
  Real dv[D], dw[A];
#pragma unroll
  for (int c = 0; c < D; ++c) dv[c] = Real(0);
#pragma unroll
  for (int k = 0; k < A; ++k) dw[k] = Real(0);
  for (int k = a.cl_start[c0]; k < a.cl_start[c0 + 1]; ++k) {
    const Real* sm = refl_row<D>(a, a.cl_members[k]);
#pragma unroll
    for (int c = 0; c < D; ++c) dv[c] = dv[c] + sm[7 * D + 2 + c];
#pragma unroll
    for (int kk = 0; kk < A; ++kk) dw[kk] = dw[kk] + sm[8 * D + 2 + kk];
  }
  if (a.fixed[i]) {
#pragma unroll
    for (int c = 0; c < D; ++c) dv[c] = Real(0);
#pragma unroll
    for (int k = 0; k < A; ++k) dw[k] = Real(0);
  }
  const Real dt_rem = (Real(1) - refl_row<D>(a, i)[7 * D + 1]) * a.dt;
  Vec<D, Real> dvv;
  from_arr(dvv, dv);
  Ang<D, Real> dwv;
  if constexpr (D == 3)
    from_arr(dwv, dw);
  else
    dwv = Ang1T<Real>{dw[0]};
  st(a.vel_out, i, ld<D>(a.vel_in, i) + dvv);
  sta(a.ang_out, i, lda<D>(a.ang_in, i) + dwv);
  Vec<3, Real> th;
  if constexpr (D == 3)
    th = dwv * dt_rem;
  else
    th = Vec<3, Real>{Real(0), Real(0), dw[0] * dt_rem};
  stq(a.q, i, q_unit(q_mul(q_from_small_rotvec(th), ldq(a.q, i))));
  st(a.pos_out, i, ld<D>(a.pos_in, i) + dvv * dt_rem);
}

// ---------------------------------------------------------------------------
// dispatch: every collider source and law in float; in double (the float64
// form of the plain step) the laws without history.
// ---------------------------------------------------------------------------
template <class Real>
constexpr bool kIsF32 = std::is_same<Real, float>::value;

#define JDEM_DISPATCH_BODY(KERNEL)                                                               \
  template <int D, int LAW>                                                                      \
  static hipError_t by_src(const StepArgsT<Real>& a, int src, bool per, dim3 g, dim3 b, hipStream_t s) { \
    switch (src) {                                                                               \
      case SRC_NAIVE:                                                                            \
        if (per)                                                                                 \
          hipLaunchKernelGGL((KERNEL<D, LAW, SRC_NAIVE, true, Real>), g, b, 0, s, a);            \
        else                                                                                     \
          hipLaunchKernelGGL((KERNEL<D, LAW, SRC_NAIVE, false, Real>), g, b, 0, s, a);           \
        break;                                                                                   \
      case SRC_CELLS:                                                                            \
        if (per)                                                                                 \
          hipLaunchKernelGGL((KERNEL<D, LAW, SRC_CELLS, true, Real>), g, b, 0, s, a);            \
        else                                                                                     \
          hipLaunchKernelGGL((KERNEL<D, LAW, SRC_CELLS, false, Real>), g, b, 0, s, a);           \
        break;                                                                                   \
      case SRC_NLIST:                                                                            \
        if (per)                                                                                 \
          hipLaunchKernelGGL((KERNEL<D, LAW, SRC_NLIST, true, Real>), g, b, 0, s, a);            \
        else                                                                                     \
          hipLaunchKernelGGL((KERNEL<D, LAW, SRC_NLIST, false, Real>), g, b, 0, s, a);           \
        break;                                                                                   \
      default:                                                                                   \
        return hipErrorInvalidValue;                                                             \
    }                                                                                            \
    return hipGetLastError();                                                                    \
  }                                                                                              \
  template <int D>                                                                               \
  static hipError_t by_law(const StepArgsT<Real>& a, int law, int src, bool per, dim3 g, dim3 b, hipStream_t s) { \
    switch (law) {                                                                               \
      case JDEM_LAW_SPRING:                                                                      \
        return by_src<D, JDEM_LAW_SPRING>(a, src, per, g, b, s);                                 \
      case JDEM_LAW_HERTZ:                                                                       \
        return by_src<D, JDEM_LAW_HERTZ>(a, src, per, g, b, s);                                  \
      case JDEM_LAW_CUNDALL_STRACK:                                                              \
        return by_src<D, JDEM_LAW_CUNDALL_STRACK>(a, src, per, g, b, s);                         \
      case kLawRouted:                                                                           \
        return by_src<D, kLawRouted>(a, src, per, g, b, s);                                      \
      case JDEM_LAW_SPRING_FRICTION:                                                             \
        if constexpr (kIsF32<Real>) return by_src<D, JDEM_LAW_SPRING_FRICTION>(a, src, per, g, b, s); \
        return hipErrorInvalidValue;                                                             \
      case JDEM_LAW_HERTZ_FRICTION:                                                              \
        if constexpr (kIsF32<Real>) return by_src<D, JDEM_LAW_HERTZ_FRICTION>(a, src, per, g, b, s); \
        return hipErrorInvalidValue;                                                             \
      case kLawRoutedHist:                                                                       \
        if constexpr (kIsF32<Real>) return by_src<D, kLawRoutedHist>(a, src, per, g, b, s);     \
        return hipErrorInvalidValue;                                                             \
      default:                                                                                   \
        return hipErrorInvalidValue;                                                             \
    }                                                                                            \
  }                                                                                              \
  static hipError_t run(const StepArgsT<Real>& a, int dim, int law, int src, hipStream_t s) {    \
    if (a.n == 0) return hipSuccess;                                                             \
    dim3 g(nblocks(a.n)), b(kBlock);                                                             \
    const bool per = a.dom.kind == JDEM_DOMAIN_PERIODIC;                                         \
    return dim == 3 ? by_law<3>(a, law, src, per, g, b, s) : by_law<2>(a, law, src, per, g, b, s); \
  }

template <class Real>
struct StepDispatch {
  JDEM_DISPATCH_BODY(k_step)
};
template <class Real>
struct EnergyDispatch {
  JDEM_DISPATCH_BODY(k_energy)
};
template <class Real>
struct VirialDispatch {
  JDEM_DISPATCH_BODY(k_virial)
};

template <class Real>
hipError_t kinetic_run(const StepArgsT<Real>& a, int dim, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (dim == 3)
    hipLaunchKernelGGL((k_kinetic<3, Real>), dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_kinetic<2, Real>), dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_step(const StepArgs& a, int dim, int law, int src, hipStream_t s) {
  return StepDispatch<float>::run(a, dim, law, src, s);
}
hipError_t launch_energy(const StepArgs& a, int dim, int law, int src, hipStream_t s) {
  return EnergyDispatch<float>::run(a, dim, law, src, s);
}
hipError_t launch_virial(const StepArgs& a, int dim, int law, int src, hipStream_t s) {
  return VirialDispatch<float>::run(a, dim, law, src, s);
}
hipError_t launch_kinetic(const StepArgs& a, int dim, hipStream_t s) { return kinetic_run(a, dim, s); }
hipError_t launch_step(const StepArgs64& a, int dim, int law, int src, hipStream_t s) {
  return StepDispatch<double>::run(a, dim, law, src, s);
}
hipError_t launch_energy(const StepArgs64& a, int dim, int law, int src, hipStream_t s) {
  return EnergyDispatch<double>::run(a, dim, law, src, s);
}
hipError_t launch_kinetic(const StepArgs64& a, int dim, hipStream_t s) { return kinetic_run(a, dim, s); }
template <class Real>
hipError_t clump_lab_run(const StepArgsT<Real>& a, int dim, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (dim == 3)
    hipLaunchKernelGGL((k_clump_lab<3, Real>), dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_clump_lab<2, Real>), dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}
template <class Real>
hipError_t clump_forces_run(const StepArgsT<Real>& a, int dim, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (dim == 3) {
    hipLaunchKernelGGL((k_clump_total<3, Real>), dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL((k_clump_sum<3, Real>), dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_clump_total<2, Real>), dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL((k_clump_sum<2, Real>), dim3(nblocks(a.n)), dim3(kBlock), 0, s, a);
  }
  return hipGetLastError();
}
template <class Real>
hipError_t clump_reflect_run(const StepArgsT<Real>& a, int dim, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const dim3 g(nblocks(a.n)), b(kBlock);
  if (dim == 3) {
    hipLaunchKernelGGL((k_crefl_over<3, Real>), g, b, 0, s, a);
    hipLaunchKernelGGL((k_crefl_alpha<3, Real>), g, b, 0, s, a);
    hipLaunchKernelGGL((k_crefl_impulse<3, Real>), g, b, 0, s, a);
    hipLaunchKernelGGL((k_crefl_apply<3, Real>), g, b, 0, s, a);
  } else {
    hipLaunchKernelGGL((k_crefl_over<2, Real>), g, b, 0, s, a);
    hipLaunchKernelGGL((k_crefl_alpha<2, Real>), g, b, 0, s, a);
    hipLaunchKernelGGL((k_crefl_impulse<2, Real>), g, b, 0, s, a);
    hipLaunchKernelGGL((k_crefl_apply<2, Real>), g, b, 0, s, a);
  }
  return hipGetLastError();
}
hipError_t launch_clump_reflect(const StepArgs& a, int dim, hipStream_t s) { return clump_reflect_run(a, dim, s); }
hipError_t launch_clump_reflect(const StepArgs64& a, int dim, hipStream_t s) { return clump_reflect_run(a, dim, s); }
hipError_t launch_clump_lab(const StepArgs& a, int dim, hipStream_t s) { return clump_lab_run(a, dim, s); }
hipError_t launch_clump_lab(const StepArgs64& a, int dim, hipStream_t s) { return clump_lab_run(a, dim, s); }
hipError_t launch_clump_forces(const StepArgs& a, int dim, hipStream_t s) { return clump_forces_run(a, dim, s); }
hipError_t launch_clump_forces(const StepArgs64& a, int dim, hipStream_t s) { return clump_forces_run(a, dim, s); }

hipError_t launch_stage(const StepArgs& a, int dim, int stage, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (dim == 3)
    hipLaunchKernelGGL(k_stage<3>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a, stage);
  else
    hipLaunchKernelGGL(k_stage<2>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, a, stage);
  return hipGetLastError();
}

}  // namespace jdem
