// Timed CPU/OpenMP baseline -- TEST/MEASUREMENT INFRASTRUCTURE ONLY.
//
// Restates the reference's CPU step for compute_type=mp-lists, order=morton
// (app/simulation_configuration.h:297-387) with its stage structure kept: cell
// ids in an OpenMP loop, a serial std::sort of particle ids by cell id, a serial
// SoA gather of every particle array, serial cell ranges over the whole table,
// multipass CSR lists (OpenMP count, serial scan, OpenMP fill; self included as
// on the CPU), OpenMP density, serial EOS, OpenMP forces, serial integrator.
// compute_type=sorted (app/simulation_configuration.h:497-533) is timed too:
// the same search, no lists, density and forces over the 3x3 cell stencil
// (GenericForceComputation.tpp:34-253). Built with the reference's -O3
// -ffast-math flags and -fopenmp. Only bench.py's cpu_baseline leg runs it.
#include <omp.h>

#include <cstring>

#include "sph_oracle_impl.hpp"

using namespace orc;

namespace {

struct Soa {  // Particles (core/container/SoAParticle.h:6-23)
  std::vector<float> mass, density, rx, ry, rz, vx, vy, vz, fx, fy, fz;
  std::vector<uint32_t> particle_id, cell_id;
  void resize(int64_t n, float m) {
    mass.assign(n, m);
    for (auto* v : {&density, &rx, &ry, &rz, &vx, &vy, &vz, &fx, &fy, &fz}) v->assign(n, 0.f);
    particle_id.resize(n);
    cell_id.resize(n);
    for (int64_t i = 0; i < n; ++i) particle_id[i] = (uint32_t)i;
  }
};

struct Timers {
  double t[7] = {0, 0, 0, 0, 0, 0, 0};
};
enum { T_INTEGRATION, T_SEARCH, T_LIST, T_DENSITY, T_PRESSURE, T_FORCE, T_MAIN };

}  // namespace

// compute_type 0: mp-lists (CSR lists, then density and forces over them);
// 1: sorted (density and forces straight over the cell stencil).
extern "C" int orc_baseline_run2(const orc_params* p, int64_t n, const float* rx0,
                                 const float* ry0, const float* vx0, const float* vy0,
                                 int32_t steps, int32_t compute_type, double* stage_seconds,
                                 double* checksum) {
  orc_grid g;
  if (make_grid(p, &g) != 0) return -1;
  Consts<float> c = make_consts<float>(p, g);
  Soa s;
  s.resize(n, (float)p->mass);
  std::memcpy(s.rx.data(), rx0, n * sizeof(float));
  std::memcpy(s.ry.data(), ry0, n * sizeof(float));
  std::memcpy(s.vx.data(), vx0, n * sizeof(float));
  std::memcpy(s.vy.data(), vy0, n * sizeof(float));
  std::vector<int32_t> cstart(g.n_keys, -1), cend(g.n_keys, -1);
  std::vector<int32_t> nnb(n), list;
  std::vector<int64_t> off(n + 1);
  std::vector<float> pres(n), pr(n);
  Timers tm;

  auto find_neighbors = [&]() {
    double t0 = omp_get_wtime();
    // SetParticleCellIds (GenericNeighborSearch.h:200-232)
#pragma omp parallel for schedule(runtime)
    for (int64_t i = 0; i < n; ++i)
      s.cell_id[i] = (uint32_t)make_key(g, cell_coord<float>(s.rx[i], c.lo_x, c.cs_x, g.nx),
                                        cell_coord<float>(s.ry[i], c.lo_y, c.cs_y, g.ny));
    // SortParticleIds: serial, unstable (GenericNeighborSearch.h:69-76)
    std::sort(s.particle_id.begin(), s.particle_id.end(),
              [&](uint32_t a, uint32_t b) { return s.cell_id[a] < s.cell_id[b]; });
    // SortContainer: serial gather of every array (GenericNeighborSearch.h:110-191)
    Soa o;
    o.density.resize(n); o.rx.resize(n); o.ry.resize(n); o.rz.resize(n);
    o.vx.resize(n); o.vy.resize(n); o.vz.resize(n); o.fx.resize(n); o.fy.resize(n);
    o.fz.resize(n); o.cell_id.resize(n);
    for (int64_t i = 0; i < n; ++i) {
      uint32_t k = s.particle_id[i];
      o.density[i] = s.density[k];
      o.rx[i] = s.rx[k]; o.ry[i] = s.ry[k]; o.rz[i] = s.rz[k];
      o.vx[i] = s.vx[k]; o.vy[i] = s.vy[k]; o.vz[i] = s.vz[k];
      o.fx[i] = s.fx[k]; o.fy[i] = s.fy[k]; o.fz[i] = s.fz[k];
      o.cell_id[i] = s.cell_id[k];
      s.particle_id[i] = (uint32_t)i;
    }
    s.density = o.density; s.rx = o.rx; s.ry = o.ry; s.rz = o.rz;
    s.vx = o.vx; s.vy = o.vy; s.vz = o.vz; s.fx = o.fx; s.fy = o.fy; s.fz = o.fz;
    s.cell_id = o.cell_id;
    // SetCellRanges: serial refill of the whole table + boundary loop (:92-108)
    std::fill(cstart.begin(), cstart.end(), -1);
    std::fill(cend.begin(), cend.end(), -1);
    for (int64_t i = 0; i < n; ++i) {
      uint32_t k = s.cell_id[i];
      if (i == 0 || k != s.cell_id[i - 1]) cstart[k] = (int32_t)i;
      if (i == n - 1 || k != s.cell_id[i + 1]) cend[k] = (int32_t)(i + 1);
    }
    tm.t[T_SEARCH] += omp_get_wtime() - t0;
  };

  auto build_lists = [&]() {  // MultipassNeighborLists.h:14-181
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(runtime)
    for (int64_t i = 0; i < n; ++i) {
      int32_t k = 0;
      for_each_neighbor<float>(g, c, s.rx.data(), s.ry.data(), cstart.data(), cend.data(), i,
                               true, [&](int32_t, float, float, float) { ++k; });
      nnb[i] = k;
    }
    off[0] = 0;
    for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + nnb[i];
    list.resize(off[n]);
#pragma omp parallel for schedule(runtime)
    for (int64_t i = 0; i < n; ++i) {
      int64_t o = off[i];
      for_each_neighbor<float>(g, c, s.rx.data(), s.ry.data(), cstart.data(), cend.data(), i,
                               true, [&](int32_t j, float, float, float) { list[o++] = j; });
    }
    tm.t[T_LIST] += omp_get_wtime() - t0;
  };

  // GenericForceComputation<MortonOrder> (GenericForceComputation.tpp:34-253):
  // the stencil itself, self included in the density, excluded in the forces
  auto compute_sorted = [&]() {
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(runtime)
    for (int64_t i = 0; i < n; ++i) {
      float acc = 0.f;
      for_each_neighbor<float>(g, c, s.rx.data(), s.ry.data(), cstart.data(), cend.data(), i,
                               true, [&](int32_t, float, float, float r2) {
                                 acc += density_term<float>(c, r2);
                               });
      s.density[i] = c.mc_rho * acc;
    }
    double t1 = omp_get_wtime();
    tm.t[T_DENSITY] += t1 - t0;
    for (int64_t i = 0; i < n; ++i) {
      float rho, p_, pr_;
      density_one<float>(c, s.density[i] / c.mc_rho, rho, p_, pr_);
      pres[i] = p_;
      pr[i] = pr_;
    }
    double t2 = omp_get_wtime();
    tm.t[T_PRESSURE] += t2 - t1;
#pragma omp parallel for schedule(runtime)
    for (int64_t i = 0; i < n; ++i) {
      Accum<float> a;
      for_each_neighbor<float>(g, c, s.rx.data(), s.ry.data(), cstart.data(), cend.data(), i,
                               false, [&](int32_t j, float dx, float dy, float r2) {
                                 force_pair<float>(c, dx, dy, r2, s.density[i], s.density[j],
                                                   pr[i], pr[j], s.vx[i], s.vy[i], s.vx[j],
                                                   s.vy[j], a);
                               });
      force_finish<float>(c, a, s.fx[i], s.fy[i]);
    }
    tm.t[T_FORCE] += omp_get_wtime() - t2;
  };

  auto compute = [&]() {  // MPListsForceComputation.cpp:3-148
    double t0 = omp_get_wtime();
#pragma omp parallel for schedule(runtime)
    for (int64_t i = 0; i < n; ++i) {
      float acc = 0.f;
      for (int64_t k = off[i]; k < off[i + 1]; ++k) {
        int32_t j = list[k];
        float dx, dy;
        separation(c, s.rx[i], s.ry[i], s.rx[j], s.ry[j], dx, dy);
        acc += density_term<float>(c, dx * dx + dy * dy);
      }
      s.density[i] = c.mc_rho * acc;
    }
    double t1 = omp_get_wtime();
    tm.t[T_DENSITY] += t1 - t0;
    for (int64_t i = 0; i < n; ++i) {
      float rho, p_, pr_;
      density_one<float>(c, s.density[i] / c.mc_rho, rho, p_, pr_);
      pres[i] = p_;
      pr[i] = pr_;
    }
    double t2 = omp_get_wtime();
    tm.t[T_PRESSURE] += t2 - t1;
#pragma omp parallel for schedule(runtime)
    for (int64_t i = 0; i < n; ++i) {
      Accum<float> a;
      for (int64_t k = off[i]; k < off[i + 1]; ++k) {
        int32_t j = list[k];
        if (j == i) continue;
        float dx, dy;
        separation(c, s.rx[i], s.ry[i], s.rx[j], s.ry[j], dx, dy);
        force_pair<float>(c, dx, dy, dx * dx + dy * dy, s.density[i], s.density[j], pr[i], pr[j],
                          s.vx[i], s.vy[i], s.vx[j], s.vy[j], a);
      }
      force_finish<float>(c, a, s.fx[i], s.fy[i]);
    }
    tm.t[T_FORCE] += omp_get_wtime() - t2;
  };

  auto evaluate = [&]() {
    if (compute_type == 1) {
      compute_sorted();
    } else {
      build_lists();
      compute();
    }
  };
  double t_start = omp_get_wtime();
  find_neighbors();
  evaluate();
  for (int it = 1; it < steps; ++it) {
    double t0 = omp_get_wtime();
    for (int64_t i = 0; i < n; ++i)  // Integrator::UpdateVelocities/UpdatePositions (serial)
      integrate_one<float>(c, s.rx[i], s.ry[i], s.vx[i], s.vy[i], s.fx[i], s.fy[i]);
    tm.t[T_INTEGRATION] += omp_get_wtime() - t0;
    find_neighbors();
    evaluate();
  }
  tm.t[T_MAIN] = omp_get_wtime() - t_start;
  for (int k = 0; k < 7; ++k) stage_seconds[k] = tm.t[k];
  if (checksum) {
    double cs = 0;
    for (int64_t i = 0; i < n; ++i) cs += (double)s.rx[i] + (double)s.ry[i];
    *checksum = cs;
  }
  return omp_get_max_threads();
}

extern "C" int orc_baseline_run(const orc_params* p, int64_t n, const float* rx0,
                                const float* ry0, const float* vx0, const float* vy0,
                                int32_t steps, double* stage_seconds, double* checksum) {
  return orc_baseline_run2(p, n, rx0, ry0, vx0, vy0, steps, 0, stage_seconds, checksum);
}
