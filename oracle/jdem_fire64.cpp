// jdem_fire64.cpp -- TEST INFRASTRUCTURE ONLY: float64 FIRE minimisation of
// harmonic spheres in a periodic cube on the CPU (C++/OpenMP), the full-FP64
// reference run against which the GPU's precision form is measured.  Nothing
// in the product (jaxdem_amd/) links or loads it.
//
// Follows the numpy oracle's FIRE (oracle/jdem_oracle.py: fire_update,
// minimize), i.e. minimizers/optimizers.py:241-318 and
// minimizers/routines.py:150-383, with the spring law of spring.py:97-108 and
// 138-147 (F = k delta rij / d, E = k delta^2 / 2), all in float64, and the
// periodic minimum image of periodic.py:39-79.  Pairs come from a Verlet list
// (cutoff R_ij + skin) built over a cell list and rebuilt once a sphere has
// moved half the skin; its order is the cell-stencil traversal's.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

typedef struct fire64_params {
  double box;  // cube edge
  double k;    // spring constant
  double dt, alpha_init, f_inc, f_dec, f_alpha, dt_max_scale, dt_min_scale;
  int n_min, n_bad_max;
  int64_t max_steps;
  double pe_tol, pe_diff_tol, force_tol;
} fire64_params;

typedef struct fire64_result {
  int64_t steps;
  double pe;        // potential energy per particle
  double max_grad;  // max |F| component
  double virial[9]; // sum over pairs r_ij (x) F_ij
} fire64_result;

int fire64_minimize(const fire64_params* p, int64_t n, double* pos, const double* rad, fire64_result* out,
                    int threads);

}  // extern "C"

namespace {

struct Cells {
  int dims;
  double h;
  std::vector<int64_t> start, order;
};

double wrap(double r, double L) { return r - L * std::nearbyint(r / L); }

void build_cells(Cells& c, int64_t n, const double* pos, double L, double rmax) {
  c.dims = std::max(1, (int)std::floor(L / (2.0 * rmax)));
  c.h = L / c.dims;
  const int64_t nc = (int64_t)c.dims * c.dims * c.dims;
  std::vector<int64_t> key(n), cnt(nc + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    int ci[3];
    for (int a = 0; a < 3; ++a) {
      double t = pos[3 * i + a] / L;
      t -= std::floor(t);
      int v = (int)std::floor(t * c.dims);
      ci[a] = std::min(std::max(v, 0), c.dims - 1);
    }
    key[i] = ((int64_t)ci[0] * c.dims + ci[1]) * c.dims + ci[2];
    cnt[key[i] + 1]++;
  }
  for (int64_t h = 0; h < nc; ++h) cnt[h + 1] += cnt[h];
  c.start = cnt;
  c.order.assign(n, 0);
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  for (int64_t i = 0; i < n; ++i) c.order[fill[key[i]]++] = i;
}

// Verlet list: for every sphere, the spheres within R_ij + skin, in the
// order of the cell-stencil traversal (so the force sums run in that order)
struct Verlet {
  std::vector<int64_t> start, nbr;
  std::vector<double> pos0;  // positions at the build
};

void build_verlet(Verlet& v, const Cells& c, int64_t n, const double* pos, const double* rad, double L, double skin) {
  const int d = c.dims;
  std::vector<int64_t> cnt(n + 1, 0);
  auto visit = [&](int64_t i, auto&& emit) {
    int ci[3];
    for (int a = 0; a < 3; ++a) {
      double t = pos[3 * i + a] / L;
      t -= std::floor(t);
      ci[a] = std::min(std::max((int)std::floor(t * d), 0), d - 1);
    }
    // distinct neighbour cells (fewer than 3 per axis are visited once)
    int offs[3][3], no[3];
    for (int a = 0; a < 3; ++a) {
      no[a] = 0;
      for (int o = -1; o <= 1; ++o) {
        const int cc = ((ci[a] + o) % d + d) % d;
        bool seen = false;
        for (int q = 0; q < no[a]; ++q) seen |= offs[a][q] == cc;
        if (!seen) offs[a][no[a]++] = cc;
      }
    }
    for (int x = 0; x < no[0]; ++x)
      for (int y = 0; y < no[1]; ++y)
        for (int z = 0; z < no[2]; ++z) {
          const int64_t h = ((int64_t)offs[0][x] * d + offs[1][y]) * d + offs[2][z];
          for (int64_t q = c.start[h]; q < c.start[h + 1]; ++q) {
            const int64_t j = c.order[q];
            if (j == i) continue;
            double r[3];
            for (int a = 0; a < 3; ++a) r[a] = wrap(pos[3 * i + a] - pos[3 * j + a], L);
            const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            const double R = rad[i] + rad[j] + skin;
            if (d2 < R * R) emit(j);
          }
        }
  };
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int64_t m = 0;
    visit(i, [&](int64_t) { ++m; });
    cnt[i + 1] = m;
  }
  for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  v.start = cnt;
  v.nbr.assign(cnt[n], 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int64_t w = v.start[i];
    visit(i, [&](int64_t j) { v.nbr[w++] = j; });
  }
  v.pos0.assign(pos, pos + 3 * n);
}

// forces, energy per particle (half of each pair), virial
void evaluate(const Verlet& v, int64_t n, const double* pos, const double* rad, double L, double k, double* F,
              double* E, double* W9) {
  double vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma omp parallel
  {
    double lv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double f[3] = {0, 0, 0}, e = 0.0;
      for (int64_t q = v.start[i]; q < v.start[i + 1]; ++q) {
        const int64_t j = v.nbr[q];
        double r[3];
        for (int a = 0; a < 3; ++a) r[a] = wrap(pos[3 * i + a] - pos[3 * j + a], L);
        const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        const double R = rad[i] + rad[j];
        if (d2 >= R * R || d2 == 0.0) continue;
        const double dist = std::sqrt(d2);
        const double delta = R - dist;
        const double s = k * delta / dist;
        for (int a = 0; a < 3; ++a) f[a] += s * r[a];
        e += 0.25 * k * delta * delta;  // half of the pair energy, counted from both sides
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) lv[3 * a + b] += 0.5 * r[a] * (s * r[b]);
      }
      for (int a = 0; a < 3; ++a) F[3 * i + a] = f[a];
      E[i] = e;
    }
#pragma omp critical
    for (int q = 0; q < 9; ++q) vir[q] += lv[q];
  }
  for (int q = 0; q < 9; ++q) W9[q] = vir[q];
}

// largest displacement since the list was built
double max_disp2(const Verlet& v, int64_t n, const double* pos) {
  double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double t = pos[3 * i + a] - v.pos0[3 * i + a];
      s += t * t;
    }
    m = std::max(m, s);
  }
  return m;
}

double safe_norm3(const double* x) {
  const double nn = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  return nn <= 1e-16 ? 1e-16 : nn;
}

}  // namespace

extern "C" int fire64_minimize(const fire64_params* p, int64_t n, double* pos, const double* rad, fire64_result* out,
                               int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  const double L = p->box;
  double rmax = 0.0;
  for (int64_t i = 0; i < n; ++i) rmax = std::max(rmax, rad[i]);
  std::vector<double> F(3 * n), E(n), v(3 * n, 0.0), vold(3 * n);
  Cells c;
  Verlet vl;
  double W9[9];
  const double skin = 0.1 * rmax;
  bool built = false;
  auto eval = [&](double& pe, double& gmax) {
    // the list holds every pair within R_ij while no sphere moved skin/2
    if (!built || max_disp2(vl, n, pos) > 0.25 * skin * skin) {
      build_cells(c, n, pos, L, rmax + 0.5 * skin);
      build_verlet(vl, c, n, pos, rad, L, skin);
      built = true;
    }
    evaluate(vl, n, pos, rad, L, p->k, F.data(), E.data(), W9);
    pe = 0.0;
    gmax = 0.0;
    for (int64_t i = 0; i < n; ++i) pe += E[i];
    for (int64_t i = 0; i < 3 * n; ++i) gmax = std::max(gmax, std::fabs(F[i]));
  };
  double dt = p->dt, alpha = p->alpha_init;
  int n_good = 0, n_bad = 0;
  double pe, gmax;
  eval(pe, gmax);
  double prev = INFINITY;
  int64_t steps = 0;
  const double tiny = 2.2250738585072014e-308;
  while (steps < p->max_steps) {
    const bool conv_pe = std::fabs(pe / (double)n) <= p->pe_tol;
    const double denom = std::max(std::max(std::fabs(pe), std::fabs(prev)), tiny);
    const bool conv_rel = std::fabs(pe - prev) / denom < p->pe_diff_tol;
    if (conv_pe || conv_rel || gmax <= p->force_tol) break;
    double power = 0.0;
    for (int64_t i = 0; i < 3 * n; ++i) {
      vold[i] = v[i] + F[i] * dt / 2.0;
      power += F[i] * vold[i];
    }
    const double cand_inc = std::min(dt * p->f_inc, p->dt * p->dt_max_scale);
    const double cand_dec = std::max(dt * p->f_dec, p->dt * p->dt_min_scale);
    double new_dt, new_alpha, dt_rev, vs;
    if (power > 0.0) {
      ++n_good;
      new_dt = n_good > p->n_min ? cand_inc : dt;
      new_alpha = n_good > p->n_min ? alpha * p->f_alpha : alpha;
      n_bad = 0;
      dt_rev = 0.0;
      vs = 1.0;
    } else {
      ++n_bad;
      const bool exceeded = n_bad > p->n_bad_max;
      new_dt = exceeded ? p->dt : cand_dec;
      if (exceeded) n_bad = 0;
      new_alpha = p->alpha_init;
      n_good = 0;
      dt_rev = -new_dt;
      vs = 0.0;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double vh[3], f[3], vo[3];
      for (int a = 0; a < 3; ++a) {
        f[a] = F[3 * i + a];
        vo[a] = vold[3 * i + a];
        vh[a] = vo[a] * vs + f[a] * new_dt / 2.0;
      }
      const double vn = safe_norm3(vh), fn = safe_norm3(f);
      const double mix = fn > 1e-16 ? vn / fn * new_alpha : 0.0;
      for (int a = 0; a < 3; ++a) {
        vh[a] = (vh[a] * (1.0 - new_alpha) + f[a] * mix) * vs;
        pos[3 * i + a] += vo[a] * dt_rev / 2.0 + vh[a] * new_dt / 2.0;
        v[3 * i + a] = vh[a];
      }
    }
    dt = new_dt;
    alpha = new_alpha;
    prev = pe;
    eval(pe, gmax);
    ++steps;
  }
  out->steps = steps;
  out->pe = pe / (double)n;
  out->max_grad = gmax;
  for (int q = 0; q < 9; ++q) out->virial[q] = W9[q];
  return 0;
}
