// Strict-IEEE build of the CPU oracle (FP32 canonical restatement + FP64 oracle).
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker; never on a product path.
#include <cstring>

#include "sph_oracle_impl.hpp"

using namespace orc;

extern "C" {

int orc_make_grid(const orc_params* p, orc_grid* g) { return make_grid(p, g); }

int orc_make_consts(const orc_params* p, orc_consts* o) {
  orc_grid g;
  int rc = make_grid(p, &g);
  if (rc != 0) return rc;
  Consts<float> c = make_consts<float>(p, g);
  o->h = c.h; o->h2 = c.h2; o->tol = c.tol; o->mc_rho = c.mc_rho; o->c_pres = c.c_pres;
  o->c_visc = c.c_visc; o->gx = c.gx; o->gy = c.gy; o->eos_k = c.eos_k; o->rho0 = c.rho0;
  o->inv_rho0 = c.inv_rho0; o->tait_b = c.tait_b; o->neg_alpha_c0 = c.neg_alpha_c0;
  o->eps_h2 = c.eps_h2; o->dt = c.dt; o->lo_x = c.lo_x; o->lo_y = c.lo_y; o->hi_x = c.hi_x;
  o->hi_y = c.hi_y; o->len_x = c.len_x; o->len_y = c.len_y; o->physics = c.physics;
  o->boundary = c.boundary; o->gamma = c.gamma; o->pad = 0;
  return 0;
}

uint64_t orc_morton2d(uint32_t cx, uint32_t cy) { return morton_loop(cx, cy); }
uint64_t orc_morton2d_magic(uint32_t cx, uint32_t cy) { return morton_magic(cx, cy); }
void orc_morton2d_decode(uint64_t key, uint32_t* cx, uint32_t* cy) {
  *cx = (uint32_t)compact1by1(key);
  *cy = (uint32_t)compact1by1(key >> 1);
}
int32_t orc_cell_coord(float x, float lo, float cs, int32_t n) {
  return cell_coord<float>(x, lo, cs, n);
}

void orc_keys_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                  uint64_t* keys) {
  orc_grid g;
  if (make_grid(p, &g) != 0) return;
  Consts<float> c = make_consts<float>(p, g);
  for (int64_t i = 0; i < n; ++i)
    keys[i] = make_key(g, cell_coord<float>(rx[i], c.lo_x, c.cs_x, g.nx),
                       cell_coord<float>(ry[i], c.lo_y, c.cs_y, g.ny));
}

void orc_stable_sort(int64_t n, const uint64_t* keys, uint32_t* perm, uint64_t* sorted_keys) {
  stable_sort_keys(n, keys, perm, sorted_keys);
}

void orc_cell_ranges(int64_t n, const uint64_t* sorted_keys, int64_t n_keys, int32_t* start,
                     int32_t* end) {
  cell_ranges(n, sorted_keys, n_keys, start, end);
}

int64_t orc_neighbor_csr_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                             const int32_t* start, const int32_t* end, int include_self,
                             int32_t* counts, int64_t* offsets, int32_t* indices) {
  return neighbor_csr<float>(p, n, rx, ry, start, end, include_self, counts, offsets, indices);
}
int64_t orc_neighbor_csr_f64(const orc_params* p, int64_t n, const double* rx, const double* ry,
                             const int32_t* start, const int32_t* end, int include_self,
                             int32_t* counts, int64_t* offsets, int32_t* indices) {
  return neighbor_csr<double>(p, n, rx, ry, start, end, include_self, counts, offsets, indices);
}

// All-pairs form of the neighbor predicate, the definition the reference's own
// brute-force path uses (core/alg/SoAForceComputation.cpp:93-112): j != i with
// r^2 <= h^2, minimum image for periodic boxes. Rows ascend in j.
int64_t orc_brute_neighbors_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                                int32_t* counts, int64_t* offsets, int32_t* indices) {
  orc_grid g;
  if (make_grid(p, &g) != 0) return -1;
  Consts<float> c = make_consts<float>(p, g);
  offsets[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    int32_t k = 0;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      float dx, dy;
      separation(c, rx[i], ry[i], rx[j], ry[j], dx, dy);
      float r2 = dx * dx + dy * dy;
      if (r2 > c.h2) continue;
      if (indices) indices[offsets[i] + k] = (int32_t)j;
      ++k;
    }
    counts[i] = k;
    offsets[i + 1] = offsets[i] + k;
  }
  return offsets[n];
}

void orc_density_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                     const int64_t* offsets, const int32_t* indices, float* rho, float* pres,
                     float* pr) {
  density_lists<float>(p, n, rx, ry, offsets, indices, rho, pres, pr);
}
void orc_density_f64(const orc_params* p, int64_t n, const double* rx, const double* ry,
                     const int64_t* offsets, const int32_t* indices, double* rho, double* pres,
                     double* pr) {
  density_lists<double>(p, n, rx, ry, offsets, indices, rho, pres, pr);
}
void orc_forces_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                    const float* vx, const float* vy, const float* rho, const float* pr,
                    const int64_t* offsets, const int32_t* indices, float* ax, float* ay) {
  forces_lists<float>(p, n, rx, ry, vx, vy, rho, pr, offsets, indices, ax, ay);
}
void orc_forces_f64(const orc_params* p, int64_t n, const double* rx, const double* ry,
                    const double* vx, const double* vy, const double* rho, const double* pr,
                    const int64_t* offsets, const int32_t* indices, double* ax, double* ay) {
  forces_lists<double>(p, n, rx, ry, vx, vy, rho, pr, offsets, indices, ax, ay);
}
void orc_integrate_f32(const orc_params* p, int64_t n, float* rx, float* ry, float* vx,
                       float* vy, const float* ax, const float* ay) {
  integrate_all<float>(p, n, rx, ry, vx, vy, ax, ay);
}
void orc_integrate_f64(const orc_params* p, int64_t n, double* rx, double* ry, double* vx,
                       double* vy, const double* ax, const double* ay) {
  integrate_all<double>(p, n, rx, ry, vx, vy, ax, ay);
}

// ComputeKineticEnergy / ComputeMaxVelocity / ParticlesInDomain
// (core/container/MeasuringFunctions.h:5-62), accumulated in double with Kahan
// compensation; the reference reports max |v|^2 under the name |v|_max.
void orc_measure_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                     const float* vx, const float* vy, double* out) {
  double ke = 0, kc = 0, pe = 0, pc = 0, vmax = 0;
  int64_t inside = 0;
  double m = p->mass, g = -p->gravity_y;
  float lox = (float)p->lo[0], loy = (float)p->lo[1], hix = (float)p->hi[0],
        hiy = (float)p->hi[1];
  for (int64_t i = 0; i < n; ++i) {
    double v2 = (double)vx[i] * (double)vx[i] + (double)vy[i] * (double)vy[i];
    double t = 0.5 * m * v2 - kc;
    double s = ke + t;
    kc = (s - ke) - t;
    ke = s;
    double e = m * g * ((double)ry[i] - (double)loy) - pc;
    double s2 = pe + e;
    pc = (s2 - pe) - e;
    pe = s2;
    if (v2 > vmax) vmax = v2;
    bool in = std::isfinite(rx[i]) && std::isfinite(ry[i]) && !(rx[i] < lox) &&
              !(rx[i] > hix) && !(ry[i] < loy) && !(ry[i] > hiy);
    inside += in ? 1 : 0;
  }
  out[0] = ke;
  out[1] = vmax;
  out[2] = (double)inside;
  out[3] = pe;
  out[4] = m * (double)n;
}

int orc_run_f32(const orc_params* p, int64_t n, float* rx, float* ry, float* vx, float* vy,
                float* ax, float* ay, float* rho, uint32_t* pid, int32_t steps) {
  return run<float>(p, n, rx, ry, vx, vy, ax, ay, rho, pid, steps);
}
int orc_run_f64(const orc_params* p, int64_t n, double* rx, double* ry, double* vx,
                double* vy, double* ax, double* ay, double* rho, uint32_t* pid, int32_t steps) {
  return run<double>(p, n, rx, ry, vx, vy, ax, ay, rho, pid, steps);
}

void orc_delta_encode(int32_t m, const int32_t* list, int32_t* out) {
  if (m <= 0) return;
  out[0] = list[0];  // first neighbor verbatim (Encoder.h:36-38)
  for (int32_t i = 1; i < m; ++i) out[i] = list[i] - list[i - 1] - 1;  // Encoder.h:40-44
}
void orc_delta_decode(int32_t m, const int32_t* d, int32_t* out) {
  if (m <= 0) return;
  out[0] = d[0];
  for (int32_t i = 1; i < m; ++i) out[i] = out[i - 1] + d[i] + 1;
}

}  // extern "C"
