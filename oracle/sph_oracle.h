/*
 * CPU oracle for the 2D SPH step -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, on the CPU, the algorithm of the reference's hot path
 * (jose-pinzon0407/sph-simulations, core/):
 *   Morton / row-major cell keys, stable sort by cell key, cell start/end ranges,
 *   self-excluded CSR neighbor lists in stencil order, summation density, equation
 *   of state, pressure + viscosity accelerations, symplectic-Euler integration with
 *   clamp-and-reflect walls, FP64 step diagnostics, and the delta neighbor encoder.
 *
 * It is built twice from the same source: once in FP32 with strict IEEE arithmetic
 * (no FP contraction, no fast-math) so integer outputs (keys, permutation, ranges,
 * neighbor sets) are the canonical bit-exact definition, and once in FP64 as the
 * accuracy oracle for density / acceleration / positions.
 *
 * Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker or the timed CPU baseline -- never as a product path.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical and numerical parameters; field meaning matches sph_params of the
 * product C-ABI (include/sph_amd.h) so one Python dict can fill both. */
typedef struct orc_params {
  double lo[2];
  double hi[2];
  double h;          /* kernel support radius */
  double cell_size;  /* <= 0 -> h */
  double dt;
  double mass;       /* uniform particle mass */
  double rho0;       /* rest density */
  double gravity_y;
  double eos_k;      /* reference physics: p = eos_k (rho - rho0) */
  double mu;         /* reference physics: viscosity */
  double c0;         /* WCSPH: numerical sound speed */
  double alpha;      /* WCSPH: Monaghan artificial viscosity */
  double eps_visc;   /* WCSPH: mu_ij regulariser (times h^2) */
  int32_t gamma;     /* WCSPH: integer Tait exponent */
  int32_t physics;   /* 0 reference (Mueller-style, 3D constants), 1 WCSPH (2D) */
  int32_t boundary;  /* 0 clamp-and-reflect walls, 1 periodic */
  int32_t order;     /* 0 Morton, 1 row-major */
  int32_t ell_width; /* unused by the oracle */
  int32_t reserved;
} orc_params;

typedef struct orc_grid {
  float cs_x, cs_y;     /* cell edge per axis */
  float lo_x, lo_y;
  float len_x, len_y;
  int32_t nx, ny;       /* logical cells per axis */
  int32_t nxp, nyp;     /* padded extents (power of two for Morton) */
  int64_t n_keys;       /* size of the cell table = max key + 1 */
  int32_t key_bits;     /* significant key bits */
  int32_t order;
} orc_grid;

/* Constants derived from orc_params, evaluated in double then rounded to float. */
typedef struct orc_consts {
  float h, h2, tol;
  float mc_rho;      /* mass * density-kernel constant */
  float c_pres;      /* pressure-gradient prefactor (see orc_forces) */
  float c_visc;      /* reference viscosity prefactor 45 mu / (pi h^6) */
  float gx, gy;
  float eos_k, rho0, inv_rho0, tait_b;
  float neg_alpha_c0, eps_h2;
  float dt;
  float lo_x, lo_y, hi_x, hi_y;
  float len_x, len_y;
  int32_t physics, boundary, gamma, pad;
} orc_consts;

int orc_make_grid(const orc_params* p, orc_grid* g);
int orc_make_consts(const orc_params* p, orc_consts* c);

uint64_t orc_morton2d(uint32_t cx, uint32_t cy);          /* bit loop */
uint64_t orc_morton2d_magic(uint32_t cx, uint32_t cy);    /* magic-bits form */
void orc_morton2d_decode(uint64_t key, uint32_t* cx, uint32_t* cy);
int32_t orc_cell_coord(float x, float lo, float cs, int32_t n);

void orc_keys_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                  uint64_t* keys);
void orc_stable_sort(int64_t n, const uint64_t* keys, uint32_t* perm, uint64_t* sorted_keys);
void orc_cell_ranges(int64_t n, const uint64_t* sorted_keys, int64_t n_keys,
                     int32_t* start, int32_t* end);

/* CSR neighbor lists over particles already in sorted order. Returns the total
 * number of entries; fills counts[n], offsets[n+1] and (if indices != NULL)
 * indices[total]. Entries are in stencil order (x offset outer, y offset inner,
 * ascending index within a cell). include_self adds i to its own list. */
int64_t orc_neighbor_csr_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                             const int32_t* start, const int32_t* end, int include_self,
                             int32_t* counts, int64_t* offsets, int32_t* indices);
int64_t orc_neighbor_csr_f64(const orc_params* p, int64_t n, const double* rx, const double* ry,
                             const int32_t* start, const int32_t* end, int include_self,
                             int32_t* counts, int64_t* offsets, int32_t* indices);
/* O(N^2) all-pairs neighbor count with the same predicate (independent of the grid). */
int64_t orc_brute_neighbors_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                                int32_t* counts, int64_t* offsets, int32_t* indices);

/* Density, pressure and pr = p / rho^2 from self-excluded CSR lists (self term first). */
void orc_density_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                     const int64_t* offsets, const int32_t* indices,
                     float* rho, float* pres, float* pr);
void orc_density_f64(const orc_params* p, int64_t n, const double* rx, const double* ry,
                     const int64_t* offsets, const int32_t* indices,
                     double* rho, double* pres, double* pr);
void orc_forces_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                    const float* vx, const float* vy, const float* rho, const float* pr,
                    const int64_t* offsets, const int32_t* indices, float* ax, float* ay);
void orc_forces_f64(const orc_params* p, int64_t n, const double* rx, const double* ry,
                    const double* vx, const double* vy, const double* rho, const double* pr,
                    const int64_t* offsets, const int32_t* indices, double* ax, double* ay);
void orc_integrate_f32(const orc_params* p, int64_t n, float* rx, float* ry, float* vx,
                       float* vy, const float* ax, const float* ay);
void orc_integrate_f64(const orc_params* p, int64_t n, double* rx, double* ry, double* vx,
                       double* vy, const double* ax, const double* ay);

/* FP64 diagnostics: out[0]=kinetic energy (Kahan-compensated), out[1]=max |v|^2,
 * out[2]=particles inside the domain, out[3]=potential energy m*(-g)*(y-lo_y),
 * out[4]=total mass. */
void orc_measure_f32(const orc_params* p, int64_t n, const float* rx, const float* ry,
                     const float* vx, const float* vy, double* out);

/* Full pipeline: `steps` reference iterations on state held in caller arrays
 * (positions, velocities, accelerations, particle ids; updated in place and
 * left in the final sorted order). Iteration 0 is the pre-loop (sort + density
 * + forces); each further iteration integrates, re-sorts, and recomputes
 * density and forces -- the order of simulation_configuration.h:297-387. */
int orc_run_f32(const orc_params* p, int64_t n, float* rx, float* ry, float* vx, float* vy,
                float* ax, float* ay, float* rho, uint32_t* pid, int32_t steps);
int orc_run_f64(const orc_params* p, int64_t n, double* rx, double* ry, double* vx,
                double* vy, double* ax, double* ay, double* rho, uint32_t* pid, int32_t steps);

/* Delta encoder of an ascending list (Encoder.h:27-44) and its inverse
 * (EncodedListsForceComputation.cpp:23-33). */
void orc_delta_encode(int32_t m, const int32_t* list, int32_t* out);
void orc_delta_decode(int32_t m, const int32_t* deltas, int32_t* out);

#ifdef __cplusplus
}
#endif
#endif
