// jdem_device.h -- device arithmetic of the DEM time step for gfx950.
//
// Every formula here follows the reference file:line cited beside it with an
// explicit association of its floating-point operations.  The library is
// compiled with -ffp-contract=off and IEEE float32 division and square root,
// so each expression rounds exactly as the same expression evaluated op by op
// in numpy float32; ``rsqrt(x)`` is evaluated as ``1/sqrt(x)``.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jdem.h"

#define JD __device__ __forceinline__

namespace jdem {

// ---------------------------------------------------------------------------
// small vectors: Vec<3>, Vec<2>; angular quantities Ang<3> = Vec<3>, Ang<2>
// is a scalar (the z component), as State.ang_vel (N,1) in 2D.
// ---------------------------------------------------------------------------
template <int D>
struct Vec;
template <>
struct Vec<3> {
  float x, y, z;
};
template <>
struct Vec<2> {
  float x, y;
};
struct Ang1 {
  float z;
};
template <int D>
struct AngSel {
  using T = Vec<3>;
};
template <>
struct AngSel<2> {
  using T = Ang1;
};
template <int D>
using Ang = typename AngSel<D>::T;

JD Vec<3> operator+(Vec<3> a, Vec<3> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
JD Vec<3> operator-(Vec<3> a, Vec<3> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
JD Vec<3> operator*(Vec<3> a, float s) { return {a.x * s, a.y * s, a.z * s}; }
JD Vec<3> operator*(Vec<3> a, Vec<3> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
JD Vec<3> operator/(Vec<3> a, float s) { return {a.x / s, a.y / s, a.z / s}; }
JD Vec<3> operator-(Vec<3> a) { return {-a.x, -a.y, -a.z}; }
JD Vec<2> operator+(Vec<2> a, Vec<2> b) { return {a.x + b.x, a.y + b.y}; }
JD Vec<2> operator-(Vec<2> a, Vec<2> b) { return {a.x - b.x, a.y - b.y}; }
JD Vec<2> operator*(Vec<2> a, float s) { return {a.x * s, a.y * s}; }
JD Vec<2> operator*(Vec<2> a, Vec<2> b) { return {a.x * b.x, a.y * b.y}; }
JD Vec<2> operator/(Vec<2> a, float s) { return {a.x / s, a.y / s}; }
JD Vec<2> operator-(Vec<2> a) { return {-a.x, -a.y}; }
JD Ang1 operator+(Ang1 a, Ang1 b) { return {a.z + b.z}; }
JD Ang1 operator-(Ang1 a, Ang1 b) { return {a.z - b.z}; }
JD Ang1 operator*(Ang1 a, float s) { return {a.z * s}; }
JD Ang1 operator*(Ang1 a, Ang1 b) { return {a.z * b.z}; }
JD Ang1 operator/(Ang1 a, float s) { return {a.z / s}; }

// jnp.vecdot with left-to-right association (linalg.py:71)
JD float dot(Vec<3> a, Vec<3> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
JD float dot(Vec<2> a, Vec<2> b) { return a.x * b.x + a.y * b.y; }
JD float dot(Ang1 a, Ang1 b) { return a.z * b.z; }

// linalg.py:15-60
JD Vec<3> cross(Vec<3> a, Vec<3> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// 2D cross product as the (N,1) torque
JD Ang1 cross_t(Vec<2> a, Vec<2> b) { return {a.x * b.y - a.y * b.x}; }
JD Vec<3> cross_t(Vec<3> a, Vec<3> b) { return cross(a, b); }
// cross_3X3D_1X2D (linalg.py:186-231): w x r
JD Vec<3> cross_wr(Vec<3> w, Vec<3> r) { return cross(w, r); }
JD Vec<2> cross_wr(Ang1 w, Vec<2> r) { return {-w.z * r.y, w.z * r.x}; }

JD float rsq(float x) { return 1.0f / sqrtf(x); }

template <int D>
JD Vec<D> zero_vec();
template <>
JD Vec<3> zero_vec<3>() { return {0.f, 0.f, 0.f}; }
template <>
JD Vec<2> zero_vec<2>() { return {0.f, 0.f}; }
template <int D>
JD Ang<D> zero_ang();
template <>
JD Vec<3> zero_ang<3>() { return {0.f, 0.f, 0.f}; }
template <>
JD Ang1 zero_ang<2>() { return {0.f}; }

template <int D>
JD Vec<D> ld(const float* p, int64_t i);
template <>
JD Vec<3> ld<3>(const float* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
template <>
JD Vec<2> ld<2>(const float* p, int64_t i) { return {p[2 * i], p[2 * i + 1]}; }
JD void st(float* p, int64_t i, Vec<3> v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
JD void st(float* p, int64_t i, Vec<2> v) {
  p[2 * i] = v.x;
  p[2 * i + 1] = v.y;
}
template <int D>
JD Ang<D> lda(const float* p, int64_t i);
template <>
JD Vec<3> lda<3>(const float* p, int64_t i) { return ld<3>(p, i); }
template <>
JD Ang1 lda<2>(const float* p, int64_t i) { return {p[i]}; }
JD void sta(float* p, int64_t i, Vec<3> v) { st(p, i, v); }
JD void sta(float* p, int64_t i, Ang1 v) { p[i] = v.z; }

// ---------------------------------------------------------------------------
// quaternions [w, x, y, z] (utils/quaternion.py)
// ---------------------------------------------------------------------------
struct Quat {
  float w, x, y, z;
};
JD Vec<3> qv(Quat q) { return {q.x, q.y, q.z}; }
JD Quat ldq(const float* p, int64_t i) {
  const float4 v = reinterpret_cast<const float4*>(p)[i];
  return {v.x, v.y, v.z, v.w};
}
JD void stq(float* p, int64_t i, Quat q) { reinterpret_cast<float4*>(p)[i] = make_float4(q.w, q.x, q.y, q.z); }

// quaternion.py:347-367
JD Quat q_mul(Quat a, Quat b) {
  Vec<3> av = qv(a), bv = qv(b);
  float w = a.w * b.w - dot(av, bv);
  Vec<3> xyz = (bv * a.w + av * b.w) + cross(av, bv);
  return {w, xyz.x, xyz.y, xyz.z};
}
// quaternion.py:75-94
JD Quat q_unit(Quat q) {
  Vec<3> v = qv(q);
  float n2 = q.w * q.w + dot(v, v);
  float safe = (n2 == 0.0f) ? 1.0f : n2;
  float inv = rsq(safe);
  return {q.w * inv, q.x * inv, q.y * inv, q.z * inv};
}
// quaternion.py:150-180
JD Quat q_conj(Quat q) {
  Vec<3> v = qv(q);
  float n2 = q.w * q.w + dot(v, v);
  float safe = (n2 == 0.0f) ? 1.0f : n2;
  float inv_n2 = 1.0f / safe;
  return {q.w * inv_n2, (-q.x) * inv_n2, (-q.y) * inv_n2, (-q.z) * inv_n2};
}
// quaternion.py:193-242 (3D)
JD Vec<3> q_rotate(Quat q, Vec<3> v) {
  Vec<3> t = cross(qv(q), v) * 2.0f;
  return (v + t * q.w) + cross(qv(q), t);
}
// quaternion.py:247-268
JD Vec<3> q_rotate_back(Quat q, Vec<3> v) { return q_rotate(q_conj(q), v); }
// quaternion.py:132-141
JD Quat q_from_small_rotvec(Vec<3> r) {
  float n2 = dot(r, r);
  float ch = 1.0f - n2 / 8.0f;
  float sc = 0.5f - n2 / 48.0f;
  return {ch, r.x * sc, r.y * sc, r.z * sc};
}
JD Vec<3> to_rotvec3(Vec<3> v) { return v; }
// velocity_verlet_spiral.py:108: [0,0,1] * rotvec in 2D
JD Vec<3> to_rotvec3(Ang1 v) { return {0.0f * v.z, 0.0f * v.z, v.z}; }

// ---------------------------------------------------------------------------
// domain helpers (domains/__init__.py:128-160, periodic.py:39-79)
// ---------------------------------------------------------------------------
struct DomP {
  int kind;
  float box[3], inv_box[3], anchor[3];
  float restitution;
};

// Domain._displacement: periodic uses r - B*round(r*invB)
template <bool PER>
JD float disp1_(float a, float b, float B, float iB) {
  float r = a - b;
  if (PER) r = r - B * rintf(r * iB);
  return r;
}
// Domain.displacement: periodic uses r - B*round(r/B)
template <bool PER>
JD float disp1(float a, float b, float B) {
  float r = a - b;
  if (PER) r = r - B * rintf(r / B);
  return r;
}
// Domain.displacement for squared distances only.  When |r| <= B/4, r/B
// rounds to a value of magnitude <= 1/4, round(r/B) is +-0 and the result is
// r up to the sign of a zero, which a square does not see; the division is
// needed only for pairs across the periodic boundary.
template <bool PER>
JD float disp1_sq(float a, float b, float B) {
  float r = a - b;
  if (PER && !(fabsf(r) <= 0.25f * B)) r = r - B * rintf(r / B);
  return r;
}
template <bool PER>
JD float dist_sq(Vec<3> a, Vec<3> b, const DomP& d) {
  const Vec<3> r = {disp1_sq<PER>(a.x, b.x, d.box[0]), disp1_sq<PER>(a.y, b.y, d.box[1]),
                    disp1_sq<PER>(a.z, b.z, d.box[2])};
  return dot(r, r);
}
template <bool PER>
JD float dist_sq(Vec<2> a, Vec<2> b, const DomP& d) {
  const Vec<2> r = {disp1_sq<PER>(a.x, b.x, d.box[0]), disp1_sq<PER>(a.y, b.y, d.box[1])};
  return dot(r, r);
}
template <bool PER>
JD Vec<3> disp_(Vec<3> a, Vec<3> b, const DomP& d) {
  return {disp1_<PER>(a.x, b.x, d.box[0], d.inv_box[0]), disp1_<PER>(a.y, b.y, d.box[1], d.inv_box[1]),
          disp1_<PER>(a.z, b.z, d.box[2], d.inv_box[2])};
}
template <bool PER>
JD Vec<2> disp_(Vec<2> a, Vec<2> b, const DomP& d) {
  return {disp1_<PER>(a.x, b.x, d.box[0], d.inv_box[0]), disp1_<PER>(a.y, b.y, d.box[1], d.inv_box[1])};
}
template <bool PER>
JD Vec<3> disp(Vec<3> a, Vec<3> b, const DomP& d) {
  return {disp1<PER>(a.x, b.x, d.box[0]), disp1<PER>(a.y, b.y, d.box[1]), disp1<PER>(a.z, b.z, d.box[2])};
}
template <bool PER>
JD Vec<2> disp(Vec<2> a, Vec<2> b, const DomP& d) {
  return {disp1<PER>(a.x, b.x, d.box[0]), disp1<PER>(a.y, b.y, d.box[1])};
}

// ---------------------------------------------------------------------------
// pair laws
// ---------------------------------------------------------------------------
struct LawP {
  int kind, n_mat;
  const float *t0, *t1, *t2, *m0, *m1, *m2, *m3;
  float p0, p1;
  // routed laws: the terms of ordered species pair p are
  // route_terms[route_off[p] .. route_off[p + 1]) (device arrays)
  int n_species;
  const int32_t* route_off;
  const LawP* route_terms;
  const int32_t* species;
};

// Kernel kinds of a routed law (jdem_law kind ROUTED): without and with a
// history-carrying term.
constexpr int kLawRouted = JDEM_LAW_ROUTED;
constexpr int kLawRoutedHist = JDEM_LAW_ROUTED + 1;
// Material and species ids share the per-particle word of the resident
// layout: fixed | mat << 1 | species << 16.
constexpr int kMatMask = 0x7fff;

// Data of one side of a pair.
template <int D>
struct Side {
  Vec<D> x;
  float r;
  int mat;
  int sp;  // species (routed laws)
  Vec<D> v;
  Ang<D> w;
  float m;
};

template <int LAW>
struct LawTraits {
  static constexpr bool routed = (LAW == kLawRouted || LAW == kLawRoutedHist);
  // Does the law read velocities / masses of the partner?
  static constexpr bool kin = (LAW == JDEM_LAW_CUNDALL_STRACK || LAW == JDEM_LAW_SPRING_FRICTION ||
                               LAW == JDEM_LAW_HERTZ_FRICTION || routed);
  static constexpr bool hist = (LAW == JDEM_LAW_SPRING_FRICTION || LAW == JDEM_LAW_HERTZ_FRICTION ||
                                LAW == kLawRoutedHist);
  static constexpr bool torque = kin;
};

JD float tab(const float* t, int n_mat, int a, int b) { return t[a * n_mat + b]; }

// Result of one pair evaluation.
template <int D>
struct PairOut {
  Vec<D> F;
  Ang<D> T;
  Vec<D> xi;  // updated tangential spring (history laws)
  bool contact;
};

// Force on i from j.  rij = _displacement(x_i, x_j); d_sq = |rij|^2.
// spring.py:97-108
template <int D>
JD PairOut<D> law_spring(const LawP& L, const Side<D>& a, const Side<D>& b, Vec<D> rij, float d_sq) {
  PairOut<D> o;
  float R = a.r + b.r;
  float k = (L.n_mat == 1) ? L.t0[0] : tab(L.t0, L.n_mat, a.mat, b.mat);
  float inv = (d_sq == 0.0f) ? 0.0f : rsq(fmaxf(d_sq, 1e-16f));
  float r = d_sq * inv;
  float delta = fmaxf(0.0f, R - r);
  float mag = (k * delta) * inv;
  o.F = rij * mag;
  o.T = zero_ang<D>();
  o.contact = delta > 0.0f;
  return o;
}

// hertz.py:98-117 with pow(delta, 1.5) = delta*sqrt(delta)
template <int D>
JD PairOut<D> law_hertz(const LawP& L, const Side<D>& a, const Side<D>& b, Vec<D> rij, float d_sq) {
  PairOut<D> o;
  float kE = (L.n_mat == 1) ? L.t0[0] : tab(L.t0, L.n_mat, a.mat, b.mat);
  float R_star = (a.r * b.r) / (a.r + b.r);
  float k = kE * sqrtf(R_star);
  float inv = (d_sq == 0.0f) ? 0.0f : rsq(fmaxf(d_sq, 1e-16f));
  float r = d_sq * inv;
  Vec<D> n = rij * inv;
  float delta = fmaxf(0.0f, (a.r + b.r) - r);
  float mag = k * (delta * sqrtf(delta));
  o.F = n * mag;
  o.T = zero_ang<D>();
  o.contact = delta > 0.0f;
  return o;
}

// cundall_strack.py:123-193
template <int D>
JD PairOut<D> law_cundall(const LawP& L, const Side<D>& a, const Side<D>& b, Vec<D> rij, float d_sq) {
  PairOut<D> o;
  const int M = L.n_mat;
  float Ei = L.m0[a.mat], Ej = L.m0[b.mat];
  float Gi = L.m1[a.mat], Gj = L.m1[b.mat];
  float kn = ((((2.0f * Ei) * a.r) * Ej) * b.r) / (Ei * a.r + Ej * b.r);
  float kt = ((((2.0f * Gi) * a.r) * Gj) * b.r) / (Gi * a.r + Gj * b.r);
  float m_eff = (a.m * b.m) / (a.m + b.m);
  float beta = tab(L.t0, M, a.mat, b.mat);
  float mu_eff = fminf(L.m2[a.mat], L.m2[b.mat]);
  float gamma_n = (2.0f * beta) * sqrtf(kn * m_eff);
  float gamma_t = (2.0f * beta) * sqrtf(kt * m_eff);
  float inv = (d_sq == 0.0f) ? 0.0f : rsq(fmaxf(d_sq, 1e-16f));
  float r = d_sq * inv;
  Vec<D> n = rij * inv;
  float delta = (a.r + b.r) - r;
  float isc = (delta > 0.0f) ? 1.0f : 0.0f;
  delta = delta * isc;
  Vec<D> r_ci = n * (-a.r);
  Vec<D> r_cj = n * b.r;
  Vec<D> v_ci = a.v + cross_wr(a.w, r_ci);
  Vec<D> v_cj = b.v + cross_wr(b.w, r_cj);
  Vec<D> v_rel = v_ci - v_cj;
  float vn = dot(v_rel, n);
  Vec<D> vt_vec = v_rel - n * vn;
  float vt2 = dot(vt_vec, vt_vec);
  float tinv = (vt2 == 0.0f) ? 0.0f : rsq(fmaxf(vt2, 1e-16f));
  Vec<D> t = vt_vec * tinv;
  float vt = vt2 * tinv;
  float Fn = fmaxf(0.0f, kn * delta - gamma_n * vn) * isc;
  float Ft = gamma_t * vt;
  float Ft_max = mu_eff * Fn;
  Ft = fminf(Ft, Ft_max) * isc;
  o.F = n * Fn - t * Ft;
  Ang<D> tq = cross_t(r_ci, o.F);
  float mu_r_eff = fminf(L.m3[a.mat], L.m3[b.mat]);
  float R_eff = (a.r * b.r) / (a.r + b.r);
  Ang<D> w_rel = a.w - b.w;
  float w2 = dot(w_rel, w_rel);
  float wsafe = (w2 == 0.0f) ? 1.0f : w2;
  Ang<D> w_hat = w_rel * rsq(wsafe);
  o.T = tq - w_hat * ((mu_r_eff * R_eff) * Fn);
  o.contact = isc > 0.0f;
  return o;
}

// Tangential spring with Coulomb cap and history (no reference counterpart;
// Cundall & Strack 1979 in the form of Luding 2008).  HERTZ=true uses the
// Hertz-Mindlin stiffnesses kn*sqrt(R* d), kt*sqrt(R* d).
template <int D, bool HERTZ>
JD PairOut<D> law_friction(const LawP& L, const Side<D>& a, const Side<D>& b, Vec<D> rij, float d_sq, Vec<D> xi_old,
                           float dt) {
  PairOut<D> o;
  const int M = L.n_mat;
  float kn = (M == 1) ? L.t0[0] : tab(L.t0, M, a.mat, b.mat);
  float kt = (M == 1) ? L.t1[0] : tab(L.t1, M, a.mat, b.mat);
  float mu = (M == 1) ? L.t2[0] : tab(L.t2, M, a.mat, b.mat);
  float inv = (d_sq == 0.0f) ? 0.0f : rsq(fmaxf(d_sq, 1e-16f));
  float r = d_sq * inv;
  Vec<D> n = rij * inv;
  float delta = fmaxf(0.0f, (a.r + b.r) - r);
  o.contact = delta > 0.0f;
  if (!o.contact) {
    o.F = zero_vec<D>();
    o.T = zero_ang<D>();
    o.xi = zero_vec<D>();
    return o;
  }
  if (HERTZ) {
    float R_star = (a.r * b.r) / (a.r + b.r);
    float s = sqrtf(R_star * delta);
    kn = kn * s;
    kt = kt * s;
  }
  float m_eff = (a.m * b.m) / (a.m + b.m);
  float gamma_n = L.p0 * sqrtf(kn * m_eff);
  Vec<D> r_ci = n * (-a.r);
  Vec<D> r_cj = n * b.r;
  Vec<D> v_ci = a.v + cross_wr(a.w, r_ci);
  Vec<D> v_cj = b.v + cross_wr(b.w, r_cj);
  Vec<D> v_rel = v_ci - v_cj;
  float vn = dot(v_rel, n);
  Vec<D> vt = v_rel - n * vn;
  float Fn = fmaxf(0.0f, kn * delta - gamma_n * vn);
  Vec<D> xi = xi_old - n * dot(xi_old, n);
  xi = xi + vt * dt;
  Vec<D> Ft = xi * (-kt);
  float ft2 = dot(Ft, Ft);
  float fmax = mu * Fn;
  if (ft2 > fmax * fmax) {
    float scale = fmax * rsq(ft2);
    Ft = Ft * scale;
    if (kt > 0.0f) xi = Ft / (-kt);
  }
  o.F = n * Fn + Ft;
  o.T = cross_t(r_ci, o.F);
  o.xi = xi;
  return o;
}

// Routed laws (router.py:70-160, law_combiner.py:60-93): the terms of the
// pair's species entry, in order; the first gives the force and torque, each
// later one is added to them.  A LawCombiner entry starts with a NONE term
// (the combiner's zeros); the entry's history term, if any, gives the
// tangential spring and the contact flag.
JD int route_of(const LawP& L, int sa, int sb) { return L.n_species > 1 ? sa * L.n_species + sb : 0; }

template <int D, bool H>
JD PairOut<D> law_routed(const LawP& L, const Side<D>& a, const Side<D>& b, Vec<D> rij, float d_sq, Vec<D> xi_old,
                         float dt) {
  const int p = route_of(L, a.sp, b.sp);
  const int k0 = L.route_off[p], k1 = L.route_off[p + 1];
  PairOut<D> o;
  o.F = zero_vec<D>();
  o.T = zero_ang<D>();
  o.xi = zero_vec<D>();
  o.contact = false;
  for (int k = k0; k < k1; ++k) {
    const LawP& T = L.route_terms[k];
    PairOut<D> r;
    r.F = zero_vec<D>();
    r.T = zero_ang<D>();
    switch (T.kind) {
      case JDEM_LAW_SPRING:
        r = law_spring<D>(T, a, b, rij, d_sq);
        break;
      case JDEM_LAW_HERTZ:
        r = law_hertz<D>(T, a, b, rij, d_sq);
        break;
      case JDEM_LAW_CUNDALL_STRACK:
        r = law_cundall<D>(T, a, b, rij, d_sq);
        break;
      case JDEM_LAW_SPRING_FRICTION:
        if constexpr (H) {
          r = law_friction<D, false>(T, a, b, rij, d_sq, xi_old, dt);
          o.xi = r.xi;
          o.contact = r.contact;
        }
        break;
      case JDEM_LAW_HERTZ_FRICTION:
        if constexpr (H) {
          r = law_friction<D, true>(T, a, b, rij, d_sq, xi_old, dt);
          o.xi = r.xi;
          o.contact = r.contact;
        }
        break;
      default:
        break;
    }
    if (k == k0) {
      o.F = r.F;
      o.T = r.T;
    } else {
      o.F = o.F + r.F;
      o.T = o.T + r.T;
    }
  }
  return o;
}

template <int D, int LAW>
JD PairOut<D> law_eval(const LawP& L, const Side<D>& a, const Side<D>& b, Vec<D> rij, float d_sq, Vec<D> xi_old,
                       float dt) {
  if constexpr (LAW == JDEM_LAW_SPRING) return law_spring<D>(L, a, b, rij, d_sq);
  if constexpr (LAW == JDEM_LAW_HERTZ) return law_hertz<D>(L, a, b, rij, d_sq);
  if constexpr (LAW == JDEM_LAW_CUNDALL_STRACK) return law_cundall<D>(L, a, b, rij, d_sq);
  if constexpr (LAW == JDEM_LAW_SPRING_FRICTION) return law_friction<D, false>(L, a, b, rij, d_sq, xi_old, dt);
  if constexpr (LAW == JDEM_LAW_HERTZ_FRICTION) return law_friction<D, true>(L, a, b, rij, d_sq, xi_old, dt);
  if constexpr (LAW == kLawRouted) return law_routed<D, false>(L, a, b, rij, d_sq, xi_old, dt);
  if constexpr (LAW == kLawRoutedHist) return law_routed<D, true>(L, a, b, rij, d_sq, xi_old, dt);
  PairOut<D> o;
  o.F = zero_vec<D>();
  o.T = zero_ang<D>();
  o.contact = false;
  return o;
}

// Pair energy (spring.py:138-147, hertz.py:147-162, cundall_strack.py:220-233);
// the history laws report their elastic normal part.
template <int D, int LAW>
JD float law_energy(const LawP& L, const Side<D>& a, const Side<D>& b, float d_sq) {
  const int M = L.n_mat;
  if constexpr (LAW == JDEM_LAW_SPRING || LAW == JDEM_LAW_SPRING_FRICTION) {
    float k = (M == 1) ? L.t0[0] : tab(L.t0, M, a.mat, b.mat);
    float inv = (d_sq == 0.0f) ? 0.0f : rsq(fmaxf(d_sq, 1e-16f));
    float r = d_sq * inv;
    float s = fmaxf(0.0f, (a.r + b.r) - r);
    return (0.5f * k) * (s * s);
  }
  float r = (d_sq == 0.0f) ? 0.0f : sqrtf(fmaxf(d_sq, 1e-16f));
  float delta = (a.r + b.r) - r;
  delta = delta * ((delta > 0.0f) ? 1.0f : 0.0f);
  if constexpr (LAW == JDEM_LAW_HERTZ || LAW == JDEM_LAW_HERTZ_FRICTION) {
    float kE = (M == 1) ? L.t0[0] : tab(L.t0, M, a.mat, b.mat);
    float R_star = (a.r * b.r) / (a.r + b.r);
    float k = kE * sqrtf(R_star);
    return (0.4f * k) * ((delta * delta) * sqrtf(delta));
  }
  if constexpr (LAW == JDEM_LAW_CUNDALL_STRACK) {
    float Ei = L.m0[a.mat], Ej = L.m0[b.mat];
    float kn = ((((2.0f * Ei) * a.r) * Ej) * b.r) / (Ei * a.r + Ej * b.r);
    return (0.5f * kn) * (delta * delta);
  }
  if constexpr (LAW == kLawRouted || LAW == kLawRoutedHist) {
    // router.py:262-330 selects the entry; law_combiner.py:180-188 sums it
    const int p = route_of(L, a.sp, b.sp);
    const int k0 = L.route_off[p], k1 = L.route_off[p + 1];
    float e = 0.0f;
    for (int k = k0; k < k1; ++k) {
      const LawP& T = L.route_terms[k];
      float ek = 0.0f;
      switch (T.kind) {
        case JDEM_LAW_SPRING:
        case JDEM_LAW_SPRING_FRICTION:
          ek = law_energy<D, JDEM_LAW_SPRING>(T, a, b, d_sq);
          break;
        case JDEM_LAW_HERTZ:
        case JDEM_LAW_HERTZ_FRICTION:
          ek = law_energy<D, JDEM_LAW_HERTZ>(T, a, b, d_sq);
          break;
        case JDEM_LAW_CUNDALL_STRACK:
          ek = law_energy<D, JDEM_LAW_CUNDALL_STRACK>(T, a, b, d_sq);
          break;
        default:
          break;
      }
      e = (k == k0) ? ek : e + ek;
    }
    return e;
  }
  return 0.0f;
}

// ---------------------------------------------------------------------------
// integrators
// ---------------------------------------------------------------------------
// spiral.py:24-40
JD Vec<3> omega_dot(Vec<3> w, Vec<3> tq, Vec<3> I, Vec<3> invI) {
  float cx = w.y * (I.z * w.z) - w.z * (I.y * w.y);
  float cy = w.z * (I.x * w.x) - w.x * (I.z * w.z);
  float cz = w.x * (I.y * w.y) - w.y * (I.x * w.x);
  return (tq - Vec<3>{cx, cy, cz}) * invI;
}
JD Ang1 omega_dot(Ang1 w, Ang1 tq, Ang1 I, Ang1 invI) { return tq * invI; }

JD Vec<3> inv_of(Vec<3> I) { return {1.0f / I.x, 1.0f / I.y, 1.0f / I.z}; }
JD Ang1 inv_of(Ang1 I) { return {1.0f / I.z}; }

JD Vec<3> body_of(Quat q, Vec<3> v) { return q_rotate_back(q, v); }
JD Ang1 body_of(Quat, Ang1 v) { return v; }
JD Vec<3> lab_of(Quat q, Vec<3> v) { return q_rotate(q, v); }
JD Ang1 lab_of(Quat, Ang1 v) { return v; }

// SSPRK3 kick of velocity_verlet_spiral.py:100-106 (also 168-174)
template <class A>
JD A ssprk3_half(A w, A tq, A I, A invI, float dt_2, float mask) {
  A k1 = omega_dot(w, tq, I, invI) * dt_2;
  A k2 = omega_dot(w + k1, tq, I, invI) * dt_2;
  A k3 = omega_dot(w + (k1 + k2) * 0.25f, tq, I, invI) * dt_2;
  return w + (((k1 + k2) + k3 * 4.0f) * mask) / 6.0f;
}

// VelocityVerletSpiral.step_before_force (velocity_verlet_spiral.py:83-116)
template <int D>
JD void vvs_before(Quat& q, Ang<D>& w_lab, Ang<D> tq_lab, Ang<D> I, float dt, float mask) {
  float dt_2 = dt / 2.0f;
  Ang<D> invI = inv_of(I);
  Ang<D> w = body_of(q, w_lab);
  Ang<D> tq = body_of(q, tq_lab);
  w = ssprk3_half(w, tq, I, invI, dt_2, mask);
  Vec<3> rv = to_rotvec3(w * (2.0f * dt_2));
  Quat dq = q_from_small_rotvec(rv);
  q = q_unit(q_mul(q, dq));
  w_lab = lab_of(q, w);
}
// VelocityVerletSpiral.step_after_force (velocity_verlet_spiral.py:156-180)
template <int D>
JD void vvs_after(Quat q, Ang<D>& w_lab, Ang<D> tq_lab, Ang<D> I, float dt, float mask) {
  float dt_2 = dt / 2.0f;
  Ang<D> invI = inv_of(I);
  Ang<D> w = body_of(q, w_lab);
  Ang<D> tq = body_of(q, tq_lab);
  w = ssprk3_half(w, tq, I, invI, dt_2, mask);
  w_lab = lab_of(q, w);
}
// Spiral.step_after_force (spiral.py:60-144)
template <int D>
JD void spiral_after(Quat& q, Ang<D>& w_lab, Ang<D> tq_lab, Ang<D> I, float dt, float mask) {
  float dt_2 = dt / 2.0f;
  Ang<D> invI = inv_of(I);
  Ang<D> w = body_of(q, w_lab);
  Ang<D> tq = body_of(q, tq_lab);
  Ang<D> w_dot = omega_dot(w, tq, I, invI);
  Vec<3> rv1 = to_rotvec3(w * (2.0f * dt_2));
  Vec<3> rv2 = to_rotvec3(w_dot * ((2.0f * dt_2) * dt_2));
  Quat dq = q_mul(q_from_small_rotvec(rv1), q_from_small_rotvec(rv2));
  q = q_unit(q_mul(q, dq));
  Ang<D> k1 = w_dot * dt;
  Ang<D> k2 = omega_dot(w + k1, tq, I, invI) * dt;
  Ang<D> k3 = omega_dot(w + (k1 + k2) * 0.25f, tq, I, invI) * dt;
  w = w + (((k1 + k2) + k3 * 4.0f) * mask) / 6.0f;
  w_lab = lab_of(q, w);
}

// ---------------------------------------------------------------------------
// reflect domain (reflect.py:165-297, _toc.py:11-94), one sphere
// ---------------------------------------------------------------------------
JD float toc_alpha(float v, float acc, float delta, float ws, float dt) {
  float v0 = v - dt * acc;
  float vmid = v - (0.5f * dt) * acc;
  float v0n = v0 * ws;
  float vmidn = vmid * ws;
  float accn = acc * ws;
  float A = ((0.5f * accn) * dt) * dt;
  float B = v0n * dt;
  float dw = delta + vmidn * dt;
  float disc = B * B + (4.0f * A) * dw;
  disc = fmaxf(0.0f, disc);
  float sq = (disc > 0.0f) ? sqrtf(disc) : 0.0f;
  float alpha;
  if (B < 0.0f) {
    float den = B - sq;
    alpha = (2.0f * dw) / ((den < -1e-10f) ? den : -1.0f);
  } else {
    float twoA = 2.0f * A;
    alpha = ((-B) - sq) / ((fabsf(twoA) > 1e-10f) ? twoA : 1.0f);
  }
  // jnp.clip(alpha, 0, 1) (NaN-free inputs)
  return fminf(fmaxf(alpha, 0.0f), 1.0f);
}

JD void to_arr(Vec<3> v, float* a) {
  a[0] = v.x;
  a[1] = v.y;
  a[2] = v.z;
}
JD void to_arr(Vec<2> v, float* a) {
  a[0] = v.x;
  a[1] = v.y;
}
JD void from_arr(Vec<3>& v, const float* a) { v = {a[0], a[1], a[2]}; }
JD void from_arr(Vec<2>& v, const float* a) { v = {a[0], a[1]}; }

// Register arrays with fully unrolled component loops (no stack traffic).
template <int D>
JD void reflect_apply(const DomP& dom, Vec<D>& x, Vec<D>& v, Quat& q, Vec<D> F, float m, float rad, bool fixed,
                      float dt) {
  float xs[D], vs[D], fs[D];
  to_arr(x, xs);
  to_arr(v, vs);
  to_arr(F, fs);
  const float e = dom.restitution;
  const float inv_mass = 1.0f / m;
  float denom = inv_mass + 0.0f;
  denom = (denom == 0.0f) ? 1.0f : denom;
  float ws[D], act[D], acc[D];
  float a_min = 1.0f;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const float lo = dom.anchor[c] + rad;
    const float hi = (dom.anchor[c] + dom.box[c]) - rad;
    const float over_lo = fmaxf(0.0f, lo - xs[c]);
    const float over_hi = fmaxf(0.0f, xs[c] - hi);
    const float delta = fmaxf(over_lo, over_hi);
    ws[c] = ((over_lo > 0.0f) ? 1.0f : 0.0f) - ((over_hi > 0.0f) ? 1.0f : 0.0f);
    act[c] = fabsf(ws[c]);
    acc[c] = fs[c] * inv_mass;
    const float alpha = toc_alpha(vs[c], acc[c], delta, ws[c], dt);
    if (act[c] > 0.0f) a_min = fminf(a_min, alpha);
  }
  const float dt_factor = (a_min - 1.0f) * dt;
  const float dt_rem = (1.0f - a_min) * dt;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const float v_contact = vs[c] + dt_factor * acc[c];
    float j = ((-(1.0f + e)) * v_contact) / denom;
    const float closing = ((v_contact * ws[c]) < 0.0f) ? 1.0f : 0.0f;
    const float weight = act[c] / fmaxf(act[c], 1.0f);
    j = j * (weight * closing);
    const float dv = fixed ? 0.0f : j * inv_mass;
    vs[c] = vs[c] + dv;
    xs[c] = xs[c] + dv * dt_rem;
  }
  from_arr(v, vs);
  from_arr(x, xs);
  q = q_unit(q);
}

}  // namespace jdem
