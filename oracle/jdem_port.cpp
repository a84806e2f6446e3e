// jdem_port.cpp -- TEST INFRASTRUCTURE ONLY: C++/OpenMP port of the CPU
// oracle (oracle/jdem_oracle.py), used as the timed CPU baseline of bench.py
// and checked bit for bit against the numpy oracle by the CPU tests.  Nothing
// in the product (jaxdem_amd/) links or loads it.
//
// Scope: 3D spheres, the NeighborList collider over a cell list
// (neighbor_list.py:58-131, 594-669; cell_list.py:500-595), the spring
// (spring.py:97-108) and springfriction laws, periodic or reflective domains
// (periodic.py:75-79; reflect.py:165-297), velocity Verlet
// (velocity_verlet.py:57-94) and VV-SPIRAL (velocity_verlet_spiral.py:
// 83-180), in the step order of system.py:60-82.  Arithmetic follows the
// numpy oracle op by op; build with -ffp-contract=off and no fast-math.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

typedef struct port_params {
  float dt;
  float box[3];
  float inv_box[3];
  float anchor[3];
  int periodic;
  int reflect;
  float restitution;
  int law;  // 1 spring, 4 springfriction
  float kn, kt, mu, damp2;
  float cutoff, skin;
  int M;
  float cell_size;  // secondary cell list cell size (before inflation)
  int sr;
} port_params;

int jdem_port_steps(const port_params* p, int64_t n, float* pos, float* vel, float* angv, float* q, float* force,
                    float* torque, const float* rad, const float* mass, const float* inertia, int32_t* nl,
                    float* old_pos, float* xi, uint8_t* live, int32_t* n_build, int64_t steps, int threads);
}

namespace {

struct V3 {
  float x, y, z;
};
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 mulv(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 ld(const float* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline void st(float* p, int64_t i, V3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
struct Q {
  float w, x, y, z;
};
inline V3 qv(Q q) { return {q.x, q.y, q.z}; }
inline Q qmul(Q a, Q b) {
  V3 av = qv(a), bv = qv(b);
  float w = a.w * b.w - dot(av, bv);
  V3 x = add(add(mul(bv, a.w), mul(av, b.w)), cross(av, bv));
  return {w, x.x, x.y, x.z};
}
inline Q qunit(Q q) {
  float n2 = q.w * q.w + dot(qv(q), qv(q));
  float s = n2 == 0.0f ? 1.0f : n2;
  float inv = 1.0f / std::sqrt(s);
  return {q.w * inv, q.x * inv, q.y * inv, q.z * inv};
}
// Quaternion.conj (quaternion.py:143-162), applied by rotate_back
inline Q qconj(Q q) { return {q.w, -q.x, -q.y, -q.z}; }
inline V3 qrot(Q q, V3 v) {
  V3 t = mul(cross(qv(q), v), 2.0f);
  return add(add(v, mul(t, q.w)), cross(qv(q), t));
}
inline V3 qrotb(Q q, V3 v) { return qrot(qconj(q), v); }
inline Q small_rotvec(V3 r) {
  float n2 = dot(r, r);
  float ch = 1.0f - n2 / 8.0f;
  float sc = 0.5f - n2 / 48.0f;
  return {ch, r.x * sc, r.y * sc, r.z * sc};
}
inline V3 omega_dot(V3 w, V3 tq, V3 I, V3 iI) {
  float cx = w.y * (I.z * w.z) - w.z * (I.y * w.y);
  float cy = w.z * (I.x * w.x) - w.x * (I.z * w.z);
  float cz = w.x * (I.y * w.y) - w.y * (I.x * w.x);
  return mulv(sub(tq, V3{cx, cy, cz}), iI);
}
inline V3 ssprk3(V3 w, V3 tq, V3 I, V3 iI, float dt_2, float mask) {
  V3 k1 = mul(omega_dot(w, tq, I, iI), dt_2);
  V3 k2 = mul(omega_dot(add(w, k1), tq, I, iI), dt_2);
  V3 k3 = mul(omega_dot(add(w, mul(add(k1, k2), 0.25f)), tq, I, iI), dt_2);
  V3 s = mul(add(add(k1, k2), mul(k3, 4.0f)), mask);
  return add(w, V3{s.x / 6.0f, s.y / 6.0f, s.z / 6.0f});
}

struct Ctx {
  const port_params* p;
  int64_t n;
};

inline float disp1_(float a, float b, float B, float iB, bool per) {
  float r = a - b;
  if (per) r = r - B * std::rint(r * iB);
  return r;
}
inline float disp1(float a, float b, float B, bool per) {
  float r = a - b;
  if (per) r = r - B * std::rint(r / B);
  return r;
}

float toc_alpha(float v, float acc, float delta, float ws, float dt) {
  float v0 = v - dt * acc;
  float vmid = v - (0.5f * dt) * acc;
  float v0n = v0 * ws, vmidn = vmid * ws, accn = acc * ws;
  float A = ((0.5f * accn) * dt) * dt;
  float B = v0n * dt;
  float dw = delta + vmidn * dt;
  float disc = B * B + (4.0f * A) * dw;
  disc = std::fmax(0.0f, disc);
  float sq = disc > 0.0f ? std::sqrt(disc) : 0.0f;
  float alpha;
  if (B < 0.0f) {
    float den = B - sq;
    alpha = (2.0f * dw) / (den < -1e-10f ? den : -1.0f);
  } else {
    float twoA = 2.0f * A;
    alpha = ((-B) - sq) / (std::fabs(twoA) > 1e-10f ? twoA : 1.0f);
  }
  return std::fmin(std::fmax(alpha, 0.0f), 1.0f);
}

// Neighbour list through a cell list (cell_list.py:500-595) with the
// history remap of neighbor_list.py:29-54.
void rebuild(const port_params* p, int64_t n, const float* pos, const float* rad, int32_t* nl, float* xi,
             uint8_t* live, bool first) {
  const bool per = p->periodic != 0;
  const float lc = p->cutoff + p->skin;
  const float cut_sq = lc * lc;
  const int sr = p->sr < 1 ? 1 : p->sr;
  float cs = p->cell_size;
  float infl = lc / (float)sr;
  if (infl > cs) cs = infl;
  int dims[3];
  float cell[3];
  for (int a = 0; a < 3; ++a) {
    if (per) {
      int d = (int)std::floor(p->box[a] / cs);
      dims[a] = d < 1 ? 1 : d;
      cell[a] = p->box[a] / (float)dims[a];
    } else {
      int d = (int)std::ceil(p->box[a] / cs);
      dims[a] = d < 1 ? 1 : d;
      cell[a] = cs;
    }
  }
  const int64_t ncell = (int64_t)dims[0] * dims[1] * dims[2];
  std::vector<int32_t> h(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int c[3];
    for (int a = 0; a < 3; ++a) {
      float f;
      if (per) {
        float t = (pos[3 * i + a] - p->anchor[a]) / p->box[a];
        t = t - std::floor(t);
        f = std::floor(t * (float)dims[a]);
      } else {
        f = std::floor((pos[3 * i + a] - p->anchor[a]) / cell[a]);
      }
      f = std::fmin(std::fmax(f, 0.0f), (float)(dims[a] - 1));
      c[a] = (int)f;
    }
    h[i] = c[0] + c[1] * dims[0] + c[2] * dims[0] * dims[1];
  }
  // stable counting sort by cell (ascending id within a cell)
  std::vector<int32_t> start(ncell + 1, 0), perm(n);
  for (int64_t i = 0; i < n; ++i) start[h[i] + 1]++;
  for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
  {
    std::vector<int32_t> cur(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < n; ++i) perm[cur[h[i]]++] = (int32_t)i;
  }
  const int M = p->M;
  std::vector<int32_t> newnl((size_t)n * M);
  int hi[3];
  for (int a = 0; a < 3; ++a) hi[a] = per ? std::min(sr, dims[a] - sr - 1) : sr;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    int c[3] = {h[i] % dims[0], (h[i] / dims[0]) % dims[1], h[i] / (dims[0] * dims[1])};
    int32_t* row = &newnl[(size_t)i * M];
    int count = 0;
    V3 xi_ = ld(pos, i);
    for (int ox = -sr; ox <= hi[0]; ++ox) {
      int cx = c[0] + ox;
      if (per) {
        cx %= dims[0];
        if (cx < 0) cx += dims[0];
      } else if (cx < 0 || cx >= dims[0])
        continue;
      for (int oy = -sr; oy <= hi[1]; ++oy) {
        int cy = c[1] + oy;
        if (per) {
          cy %= dims[1];
          if (cy < 0) cy += dims[1];
        } else if (cy < 0 || cy >= dims[1])
          continue;
        for (int oz = -sr; oz <= hi[2]; ++oz) {
          int cz = c[2] + oz;
          if (per) {
            cz %= dims[2];
            if (cz < 0) cz += dims[2];
          } else if (cz < 0 || cz >= dims[2])
            continue;
          int hh = cx + cy * dims[0] + cz * dims[0] * dims[1];
          for (int k = start[hh]; k < start[hh + 1]; ++k) {
            int32_t j = perm[k];
            if (j == i) continue;
            V3 xj = ld(pos, j);
            V3 d = {disp1(xi_.x, xj.x, p->box[0], per), disp1(xi_.y, xj.y, p->box[1], per),
                    disp1(xi_.z, xj.z, p->box[2], per)};
            if (!(dot(d, d) <= cut_sq)) continue;
            if (count < M) row[count] = j;
            ++count;
          }
        }
      }
    }
    for (int t = count; t < M; ++t) row[t] = -1;
  }
  if (xi && live) {
    std::vector<float> nxi((size_t)n * M * 3, 0.0f);
    std::vector<uint8_t> nlive((size_t)n * M, 0);
    if (!first) {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) {
        for (int s = 0; s < M; ++s) {
          int32_t j = newnl[(size_t)i * M + s];
          if (j < 0) continue;
          for (int u = 0; u < M; ++u) {
            if (nl[(size_t)i * M + u] == j) {
              if (live[(size_t)i * M + u]) {
                for (int c2 = 0; c2 < 3; ++c2) nxi[((size_t)i * M + s) * 3 + c2] = xi[((size_t)i * M + u) * 3 + c2];
                nlive[(size_t)i * M + s] = 1;
              }
              break;
            }
          }
        }
      }
    }
    std::memcpy(xi, nxi.data(), nxi.size() * sizeof(float));
    std::memcpy(live, nlive.data(), nlive.size());
  }
  std::memcpy(nl, newnl.data(), newnl.size() * sizeof(int32_t));
}

}  // namespace

int jdem_port_steps(const port_params* p, int64_t n, float* pos, float* vel, float* angv, float* q, float* force,
                    float* torque, const float* rad, const float* mass, const float* inertia, int32_t* nl,
                    float* old_pos, float* xi, uint8_t* live, int32_t* n_build, int64_t steps, int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  const bool per = p->periodic != 0;
  const float dt = p->dt;
  const int M = p->M;
  const bool fric = p->law == 4;
  for (int64_t step = 0; step < steps; ++step) {
    // domain.apply + step_before_force
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      V3 x = ld(pos, i), v = ld(vel, i), F = ld(force, i), w = ld(angv, i), T = ld(torque, i), I = ld(inertia, i);
      Q qq = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
      const float m = mass[i];
      if (p->reflect) {
        float* xs = &x.x;
        float* vs = &v.x;
        const float* fs = &F.x;
        float inv_mass = 1.0f / m;
        float denom = inv_mass + 0.0f;
        denom = denom == 0.0f ? 1.0f : denom;
        float ws[3], act[3], acc[3], amin = 1.0f;
        for (int c = 0; c < 3; ++c) {
          float lo = p->anchor[c] + rad[i];
          float hi = (p->anchor[c] + p->box[c]) - rad[i];
          float olo = std::fmax(0.0f, lo - xs[c]);
          float ohi = std::fmax(0.0f, xs[c] - hi);
          float delta = std::fmax(olo, ohi);
          ws[c] = (olo > 0.0f ? 1.0f : 0.0f) - (ohi > 0.0f ? 1.0f : 0.0f);
          act[c] = std::fabs(ws[c]);
          acc[c] = fs[c] * inv_mass;
          float al = toc_alpha(vs[c], acc[c], delta, ws[c], dt);
          if (act[c] > 0.0f) amin = std::fmin(amin, al);
        }
        float dtf = (amin - 1.0f) * dt, dtr = (1.0f - amin) * dt;
        float dv[3];
        for (int c = 0; c < 3; ++c) {
          float vc = vs[c] + dtf * acc[c];
          float j = ((-(1.0f + p->restitution)) * vc) / denom;
          float closing = (vc * ws[c]) < 0.0f ? 1.0f : 0.0f;
          float weight = act[c] / std::fmax(act[c], 1.0f);
          j = j * (weight * closing);
          dv[c] = j * inv_mass;
        }
        for (int c = 0; c < 3; ++c) vs[c] = vs[c] + dv[c];
        qq = qunit(qq);
        for (int c = 0; c < 3; ++c) xs[c] = xs[c] + dv[c] * dtr;
      }
      float s = (dt * 0.5f) / m;
      v = add(v, mul(mul(F, s), 1.0f));
      x = add(x, mul(v, dt));
      float dt_2 = dt / 2.0f;
      V3 iI = {1.0f / I.x, 1.0f / I.y, 1.0f / I.z};
      V3 wb = qrotb(qq, w), tb = qrotb(qq, T);
      wb = ssprk3(wb, tb, I, iI, dt_2, 1.0f);
      Q dq = small_rotvec(mul(wb, 2.0f * dt_2));
      qq = qunit(qmul(qq, dq));
      w = qrot(qq, wb);
      st(pos, i, x);
      st(vel, i, v);
      st(angv, i, w);
      q[4 * i] = qq.w;
      q[4 * i + 1] = qq.x;
      q[4 * i + 2] = qq.y;
      q[4 * i + 3] = qq.z;
    }
    // neighbour-list check (neighbor_list.py:78-85)
    float sk = p->skin * p->skin;
    float trig = sk / 4.0f;
    int need = (*n_build == 0) ? 1 : 0;
    if (!need) {
#pragma omp parallel for schedule(static) reduction(| : need)
      for (int64_t i = 0; i < n; ++i) {
        V3 d = sub(ld(pos, i), ld(old_pos, i));
        if (dot(d, d) > trig) need |= 1;
      }
    }
    if (need) {
      rebuild(p, n, pos, rad, nl, fric ? xi : nullptr, fric ? live : nullptr, *n_build == 0);
      std::memcpy(old_pos, pos, sizeof(float) * 3 * n);
      *n_build += 1;
    }
    // compute_force + step_after_force
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      V3 xi_ = ld(pos, i), vi = ld(vel, i), wi = ld(angv, i);
      const float ri = rad[i], mi = mass[i];
      V3 F = {0.0f, 0.0f, 0.0f}, T = {0.0f, 0.0f, 0.0f};
      for (int s = 0; s < M; ++s) {
        int32_t j = nl[(size_t)i * M + s];
        if (j < 0) break;
        V3 xj = ld(pos, j);
        V3 rij = {disp1_(xi_.x, xj.x, p->box[0], p->inv_box[0], per),
                  disp1_(xi_.y, xj.y, p->box[1], p->inv_box[1], per),
                  disp1_(xi_.z, xj.z, p->box[2], p->inv_box[2], per)};
        float d_sq = dot(rij, rij);
        float R = ri + rad[j];
        bool lv = fric && live[(size_t)i * M + s];
        if (fric) live[(size_t)i * M + s] = 0;
        if (d_sq > (R * R) * 1.0001f) continue;
        float inv = d_sq == 0.0f ? 0.0f : 1.0f / std::sqrt(std::fmax(d_sq, 1e-16f));
        float r = d_sq * inv;
        if (!fric) {
          float delta = std::fmax(0.0f, R - r);
          float mag = (p->kn * delta) * inv;
          F = add(F, mul(rij, mag));
          continue;
        }
        float delta = std::fmax(0.0f, (ri + rad[j]) - r);
        if (!(delta > 0.0f)) continue;
        V3 nn = mul(rij, inv);
        float m_eff = (mi * mass[j]) / (mi + mass[j]);
        float gamma_n = p->damp2 * std::sqrt(p->kn * m_eff);
        V3 r_ci = mul(nn, -ri), r_cj = mul(nn, rad[j]);
        V3 v_ci = add(vi, cross(wi, r_ci));
        V3 v_cj = add(ld(vel, j), cross(ld(angv, j), r_cj));
        V3 v_rel = sub(v_ci, v_cj);
        float vn = dot(v_rel, nn);
        V3 vt = sub(v_rel, mul(nn, vn));
        float Fn = std::fmax(0.0f, p->kn * delta - gamma_n * vn);
        V3 x0 = lv ? ld(xi, (int64_t)i * M + s) : V3{0.0f, 0.0f, 0.0f};
        V3 x1 = sub(x0, mul(nn, dot(x0, nn)));
        x1 = add(x1, mul(vt, dt));
        V3 Ft = mul(x1, -p->kt);
        float ft2 = dot(Ft, Ft);
        float fmax_ = p->mu * Fn;
        if (ft2 > fmax_ * fmax_) {
          float sc = fmax_ * (1.0f / std::sqrt(ft2));
          Ft = mul(Ft, sc);
          if (p->kt > 0.0f) x1 = V3{Ft.x / (-p->kt), Ft.y / (-p->kt), Ft.z / (-p->kt)};
        }
        V3 f = add(mul(nn, Fn), Ft);
        F = add(F, f);
        T = add(T, cross(r_ci, f));
        st(xi, (int64_t)i * M + s, x1);
        live[(size_t)i * M + s] = 1;
      }
      st(force, i, F);
      st(torque, i, T);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      V3 v = ld(vel, i), F = ld(force, i), w = ld(angv, i), T = ld(torque, i), I = ld(inertia, i);
      Q qq = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
      float s = (dt * 0.5f) / mass[i];
      v = add(v, mul(mul(F, s), 1.0f));
      float dt_2 = dt / 2.0f;
      V3 iI = {1.0f / I.x, 1.0f / I.y, 1.0f / I.z};
      V3 wb = qrotb(qq, w), tb = qrotb(qq, T);
      wb = ssprk3(wb, tb, I, iI, dt_2, 1.0f);
      w = qrot(qq, wb);
      st(vel, i, v);
      st(angv, i, w);
    }
  }
  return 0;
}
