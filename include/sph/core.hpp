// core.hpp -- C++ drop-in for the reference's core/ API (jose-pinzon0407/
// sph-simulations), running every hot-path stage on the GPU through the C ABI of
// libsph_amd.so (include/sph_amd.h).
//
// The class and method names follow the reference so its driver code
// (app/simulation_configuration.h:297-387) compiles against this header:
//   Domain                    core/Domain.h:6-38
//   Density/Pressure/ViscosityKernel  core/Kernel.h:9-79
//   Particles, SoAContainer   core/container/SoAParticle.h:6-23, SoAContainer.h:6-67
//   LatticeCreator            core/container/Lattice.h:4-37
//   MortonOrder, RowMajorOrder, GenericNeighborSearch<Order>
//                             core/alg/orders/*.h, core/alg/GenericNeighborSearch.h:9-90
//   GenericNeighborLists<Order>  core/alg/GenericNeighborLists.hpp:9-42
//   MultipassNeighborLists<Order> core/algorithmic/MultipassNeighborLists.h:7-39
//   Codifier                  core/encoders/Encoder.h:11-47
//   GenericForceComputation<Order>, ListsForceComputation, MPListsForceComputation
//                             core/compute/GenericForceComputation.hpp:5-29,
//                             core/compute/ListsForceComputation.h:5-66
//   Integrator                core/integrator/Integrator.h:12-27
//   ComputeKineticEnergy / ComputeMaxVelocity / ParticlesInDomain
//                             core/container/MeasuringFunctions.h:5-62 (FP64 on GPU)
//
// Semantics kept from the reference: FP32 Real; objects hold references to
// caller-owned containers and lists; after FindNeighbors the container is in
// sorted order and particle_id[i] = i; the CPU list convention (self included).
// Differences: every method runs on the GPU and mirrors its result into the
// host SoA (a synchronising copy); errors throw sph::Error instead of exiting.
// Physics defaults to the reference macros (core/ConfigVariables.hpp:9-15); call
// sph::SetPhysics() before the first GPU call to use the WCSPH parameters.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <sstream>
#include <string>
#include <vector>

#include "../sph_amd.h"

using Real = float;
using uint = unsigned int;
using Vec3d = std::array<Real, 3>;
using Vec3i = std::array<int, 3>;
using ScalarField = std::vector<Real>;

#ifndef H
#define H 1
#endif
#ifndef TOL
#define TOL 1.0e-12
#endif
#ifndef GAS_CONSTANT
#define GAS_CONSTANT 30.0
#endif
#ifndef RHO_0
#define RHO_0 25
#endif
#ifndef MU
#define MU 0.01
#endif

namespace sph {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void check(int rc, void* ctx, const char* what) {
  if (rc != SPH_OK)
    throw Error(std::string(what) + ": " + (sph_last_error(ctx) ? sph_last_error(ctx) : "?"));
}

// Physics and numerics shared by the GPU objects of a simulation. Defaults are
// the reference macros: linear EOS k = GAS_CONSTANT, rho0 = RHO_0, mu = MU,
// gravity -10, 3D kernel constants, clamp-and-reflect walls.
struct Physics {
  int physics = SPH_PHYS_REFERENCE;
  int boundary = SPH_BC_WALLS;
  double rho0 = RHO_0, eos_k = GAS_CONSTANT, mu = MU, gravity_y = -10.0;
  double c0 = 0.0, alpha = 0.0, eps_visc = 0.01;
  int gamma = 7;
  int ell_width = 64;
  int device = 0;
};

inline Physics& GlobalPhysics() {
  static Physics p;
  return p;
}
inline void SetPhysics(const Physics& p) { GlobalPhysics() = p; }

}  // namespace sph

// ------------------------------------------------------------------ Domain
struct Domain {
  Vec3d l_corner{0, 0, 0};
  Vec3d h_corner{10, 20, 0};

  Vec3d GetLength() const {
    Vec3d l;
    for (int i = 0; i < 3; ++i) l[i] = h_corner[i] - l_corner[i];
    return l;
  }
  bool IsInside(Vec3d r) const {
    for (int i = 0; i < 3; ++i) {
      if (!std::isfinite(r[i])) return false;
      if (r[i] < l_corner[i] || r[i] > h_corner[i]) return false;
    }
    return true;
  }
};

// ------------------------------------------------------------ kernels (host)
inline double SquaredNorm(const Vec3d& v) {
  return (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
}
inline double VecNorm(const Vec3d& v) { return std::sqrt(SquaredNorm(v)); }

class SmoothingFunction {
 public:
  explicit SmoothingFunction(Real r) : h{r}, h2{r * r} {}
  virtual ~SmoothingFunction() = default;
  virtual Real W(Vec3d& rij) = 0;
  virtual Vec3d GradW(Vec3d& r) = 0;
  virtual Real LaplacianW(Vec3d& r) = 0;

 protected:
  Real h, h2, val = 0;
};

// poly6, 315/(64 pi h^9) (core/Kernel.cpp:8-10)
class DensityKernel : public SmoothingFunction {
 public:
  explicit DensityKernel(Real r) : SmoothingFunction(r) {
    val = (Real)(315.0 / (64.0 * M_PI * std::pow((double)h, 9.0)));
  }
  Real W(Vec3d& rij) override {
    Real r2 = (Real)SquaredNorm(rij);
    return (h2 - r2) * (h2 - r2) * (h2 - r2) * val;
  }
  Vec3d GradW(Vec3d& rij) override {
    Real r1 = (Real)VecNorm(rij), r2 = (Real)SquaredNorm(rij);
    Real dW = (Real)(-6.0 * val * (h * h - r2) * (h * h - r2));
    return {dW / r1 * rij[0], dW / r1 * rij[1], dW / r1 * rij[2]};
  }
  Real LaplacianW(Vec3d& rij) override {
    Real r2 = (Real)SquaredNorm(rij);
    return (Real)(6.0 * val * (h * h - r2) * (7.0 * r2 - h * h));
  }
};

// spiky, 15/(pi h^6) (core/Kernel.cpp:38-40); GradW zero below TOL (Kernel.h:52-60)
class PressureKernel : public SmoothingFunction {
 public:
  explicit PressureKernel(Real r) : SmoothingFunction(r) {
    val = (Real)(15.0 / (M_PI * std::pow((double)h, 6.0)));
  }
  Real W(Vec3d& rij) override {
    Real r1 = (Real)VecNorm(rij);
    return (h - r1) * (h - r1) * (h - r1) * val;
  }
  Vec3d GradW(Vec3d& rij) override {
    Real r1 = (Real)VecNorm(rij);
    if (r1 < TOL) return {0, 0, 0};
    Real dW = (Real)(-3.0 * val * (h - r1) * (h - r1));
    return {dW / r1 * rij[0], dW / r1 * rij[1], dW / r1 * rij[2]};
  }
  Real LaplacianW(Vec3d&) override { return 0; }
};

// viscosity kernel, Laplacian 45/(pi h^6)(h - r) (core/Kernel.h:64-79, Kernel.cpp:64-90)
class ViscosityKernel : public SmoothingFunction {
 public:
  explicit ViscosityKernel(Real r) : SmoothingFunction(r) {
    val = (Real)(15.0 / (2.0 * M_PI * std::pow((double)h, 3.0)));
  }
  Real W(Vec3d& rij) override {
    Real r1 = (Real)VecNorm(rij);
    Real w = -0.5f * r1 * r1 * r1 / (h * h * h) + r1 * r1 / (h * h) + 0.5f * h / r1 - 1;
    return w * val;
  }
  Vec3d GradW(Vec3d& rij) override {
    Real r2 = (Real)SquaredNorm(rij), r1 = std::sqrt(r2);
    Real dW = val * (-1.5f * r2 / (h * h * h) + 2.0f * r1 / (h * h) - 0.5f * h / r2);
    return {dW / r1 * rij[0], dW / r1 * rij[1], dW / r1 * rij[2]};
  }
  Real LaplacianW(Vec3d& rij) override {
    Real r1 = (Real)VecNorm(rij);
    double h6 = std::pow((double)h, 6.0);
    return (Real)(45.0 / (M_PI * h6) * (h - r1));
  }
};

// ---------------------------------------------------------- container
struct Particles {
  std::vector<Real> mass;
  ScalarField density;
  std::vector<Real> rx, ry, rz, vx, vy, vz, fx, fy, fz;
  std::vector<uint> particle_id;
  std::vector<uint> cell_id;
};

namespace sph {
// GPU twin of one SoAContainer: the sph_ctx plus the grid it was built for.
struct DeviceState {
  void* ctx = nullptr;
  int64_t n_max = 0;
  double cell_size = 0;
  Domain domain;
  Real h = 0;
  Real dt = 0;
  ~DeviceState() {
    if (ctx) sph_destroy(ctx);
  }
};
}  // namespace sph

class SoAContainer {
 public:
  Particles& GetParticles() { return particles; }
  size_t Size() const { return size; }
  void Resize(size_t N) {
    size = N;
    particles.mass.resize(N, 1.0f);  // SoAContainer.cpp:7-27
    for (auto* v : {&particles.density, &particles.fx, &particles.fy, &particles.fz,
                    &particles.rx, &particles.ry, &particles.rz, &particles.vx, &particles.vy,
                    &particles.vz})
      v->resize(N);
    particles.cell_id.resize(N);
    particles.particle_id.resize(N);
  }
  Vec3d pos(size_t i) const { return {particles.rx[i], particles.ry[i], particles.rz[i]}; }
  Vec3d vel(size_t i) const { return {particles.vx[i], particles.vy[i], particles.vz[i]}; }
  Vec3d force(size_t i) const { return {particles.fx[i], particles.fy[i], particles.fz[i]}; }
  uint GetParticleIndex(size_t i) const { return particles.particle_id[i]; }
  void SetParticleIndex(size_t i, uint idx) { particles.particle_id[i] = idx; }
  uint GetParticleCellIndex(size_t i) const { return particles.cell_id[i]; }
  void SetParticleCellIndex(size_t i, uint idx) { particles.cell_id[i] = idx; }
  Real GetMass(size_t i) const { return particles.mass[i]; }
  void SetDensity(size_t i, Real rho) { particles.density[i] = rho; }
  Real GetDensity(size_t i) const { return particles.density[i]; }
  void SetForce(size_t i, Real x, Real y, Real z) {
    particles.fx[i] = x; particles.fy[i] = y; particles.fz[i] = z;
  }
  void SetPosition(size_t i, Real x, Real y, Real z) {
    particles.rx[i] = x; particles.ry[i] = y; particles.rz[i] = z;
  }
  void SetPosition(size_t i, Vec3d r) { SetPosition(i, r[0], r[1], r[2]); }
  void SetVelocity(size_t i, Real x, Real y, Real z) {
    particles.vx[i] = x; particles.vy[i] = y; particles.vz[i] = z;
  }
  void SetVelocity(size_t i, Vec3d v) { SetVelocity(i, v[0], v[1], v[2]); }
  void UpdatePosition(size_t i, Real dt) {  // host versions (SoAContainer.cpp:109-132)
    particles.rx[i] += dt * particles.vx[i];
    particles.ry[i] += dt * particles.vy[i];
  }
  void UpdateVelocity(size_t i, Real dt) {
    particles.vx[i] += dt * particles.fx[i];
    particles.vy[i] += dt * particles.fy[i];
  }

  // ---- GPU binding (not in the reference): the first GPU call creates the
  // device twin for (domain, cell size); host arrays are pushed before and
  // pulled after each GPU stage.
  sph::DeviceState& Device(const Domain& domain, Real cell_size) {
    if (!dev || dev->n_max < (int64_t)size || dev->cell_size != cell_size) {
      dev = std::make_shared<sph::DeviceState>();
      const sph::Physics& ph = sph::GlobalPhysics();
      sph_params p{};
      p.lo[0] = domain.l_corner[0]; p.lo[1] = domain.l_corner[1];
      p.hi[0] = domain.h_corner[0]; p.hi[1] = domain.h_corner[1];
      p.h = cell_size;
      p.cell_size = 0;
      p.dt = dt_hint;
      p.mass = size ? particles.mass[0] : 1.0;
      p.rho0 = ph.rho0; p.gravity_y = ph.gravity_y; p.eos_k = ph.eos_k; p.mu = ph.mu;
      p.c0 = ph.c0; p.alpha = ph.alpha; p.eps_visc = ph.eps_visc; p.gamma = ph.gamma;
      p.physics = ph.physics; p.boundary = ph.boundary; p.order = order_hint;
      p.ell_width = ph.ell_width;
      sph::check(sph_create(&p, (int64_t)std::max<size_t>(size, 1), ph.device, &dev->ctx),
                 nullptr, "sph_create");
      dev->n_max = (int64_t)std::max<size_t>(size, 1);
      dev->cell_size = cell_size;
      dev->domain = domain;
      dev->h = cell_size;
      dev->dt = (Real)dt_hint;
      device_current = false;
    }
    return *dev;
  }
  bool HasDevice() const { return (bool)dev; }
  void* DeviceContext() const { return dev ? dev->ctx : nullptr; }

  void PushToDevice() {
    if (!dev) throw sph::Error("no device twin");
    auto& P = particles;
    sph::check(sph_upload(dev->ctx, (int64_t)size, P.rx.data(), P.ry.data(), P.vx.data(),
                          P.vy.data(), nullptr),
               dev->ctx, "sph_upload");
    device_current = true;
  }
  void PullFromDevice(bool fields) {
    auto& P = particles;
    sph::check(sph_download(dev->ctx, P.rx.data(), P.ry.data(), P.vx.data(), P.vy.data(),
                            nullptr),
               dev->ctx, "sph_download");
    if (fields)
      sph::check(sph_download_fields(dev->ctx, P.fx.data(), P.fy.data(), P.density.data(),
                                     nullptr),
                 dev->ctx, "sph_download_fields");
  }
  // Host arrays were edited behind the GPU's back: push them before the next stage.
  void MarkHostModified() { device_current = false; }
  bool DeviceCurrent() const { return device_current; }
  void SetDeviceCurrent(bool v) { device_current = v; }

  // grid order and time step the device twin is created with
  int order_hint = SPH_ORDER_MORTON;
  double dt_hint = 1.0e-3;

 protected:
  Particles particles;
  size_t size = 0;
  std::shared_ptr<sph::DeviceState> dev;
  bool device_current = false;
};

// --------------------------------------------------------------- lattice
struct LatticeCreator {  // core/container/Lattice.h:4-37
  double dX = 0.1, dY = 0.1, dZ = 0.1;
  double xVol = 1.0, yVol = 1.0;
  Vec3i particles_per_direction{};
  int N = 0;
  void CreateLattice(Domain& domain, SoAContainer& particles) {
    particles_per_direction[0] = (int)std::ceil(xVol * domain.GetLength()[0] / dX) + 1;
    particles_per_direction[1] = (int)std::ceil(yVol * domain.GetLength()[1] / dY) + 1;
    particles_per_direction[2] = 1;
    N = particles_per_direction[0] * particles_per_direction[1];
    particles.Resize(N);
    for (int j = 0; j < particles_per_direction[1]; ++j)
      for (int i = 0; i < particles_per_direction[0]; ++i) {
        int idx = i + particles_per_direction[0] * j;
        particles.SetParticleIndex(idx, idx);
        particles.SetPosition(idx, (Real)(dX * i), (Real)(dY * j), 0);
      }
    particles.MarkHostModified();
  }
};

// ------------------------------------------------------------------ orders
struct MortonOrder {
  static constexpr int kOrder = SPH_ORDER_MORTON;
};
struct RowMajorOrder {
  static constexpr int kOrder = SPH_ORDER_ROWMAJOR;
};

template <class GridOrder>
class GenericNeighborSearch {
 public:
  GenericNeighborSearch(Domain& domain, Real cell_size) : domain(domain), cs(cell_size) {}

  // keys -> stable radix sort -> permutation -> cell ranges, on the GPU
  void FindNeighbors(SoAContainer& container) {
    container.order_hint = GridOrder::kOrder;
    sph::DeviceState& d = container.Device(domain, cs);
    if (!container.DeviceCurrent()) container.PushToDevice();
    sph::check(sph_find_neighbors(d.ctx), d.ctx, "sph_find_neighbors");
    container.PullFromDevice(true);
    sph_grid_info g{};
    sph::check(sph_get_grid(d.ctx, &g), d.ctx, "sph_get_grid");
    grid = g;
    size_t n = container.Size();
    std::vector<uint64_t> keys(n);
    sph::check(sph_get_sorted_keys(d.ctx, keys.data()), d.ctx, "sph_get_sorted_keys");
    auto& P = container.GetParticles();
    for (size_t i = 0; i < n; ++i) {
      P.cell_id[i] = (uint)keys[i];
      P.particle_id[i] = (uint)i;  // GenericNeighborSearch.h:172
    }
    cell_start.resize(g.n_keys);
    cell_end.resize(g.n_keys);
    sph::check(sph_get_cell_ranges(d.ctx, cell_start.data(), cell_end.data()), d.ctx,
               "sph_get_cell_ranges");
    container.SetDeviceCurrent(true);
    ctx = d.ctx;
  }
  std::vector<int>& GetStart() { return cell_start; }
  std::vector<int>& GetEnd() { return cell_end; }
  int GetCellStart(int cell) { return cell_start[cell]; }
  int GetCellEnd(int cell) { return cell_end[cell]; }
  int Flatten(const Vec3i& c) const {
    if (GridOrder::kOrder == SPH_ORDER_ROWMAJOR) return c[0] + c[1] * grid.nx;
    uint64_t z = 0;
    for (int i = 0; i < 32; ++i) {
      z |= (uint64_t)(((uint32_t)c[0] >> i) & 1u) << (2 * i);
      z |= (uint64_t)(((uint32_t)c[1] >> i) & 1u) << (2 * i + 1);
    }
    return (int)z;
  }
  Vec3i Unflatten(unsigned int index) const {
    if (GridOrder::kOrder == SPH_ORDER_ROWMAJOR)
      return {(int)(index % grid.nx), (int)((index / grid.nx) % grid.ny), 0};
    int x = 0, y = 0;
    for (int i = 0; i < 16; ++i) {
      x |= (int)((index >> (2 * i)) & 1u) << i;
      y |= (int)((index >> (2 * i + 1)) & 1u) << i;
    }
    return {x, y, 0};
  }
  int GetCellsInDirectionX() const { return GridOrder::kOrder == SPH_ORDER_MORTON ? grid.nxp : grid.nx; }
  int GetCellsInDirectionY() const { return GridOrder::kOrder == SPH_ORDER_MORTON ? grid.nyp : grid.ny; }
  int GetCellsInDirectionZ() const { return 1; }
  int GetTotalCells() const { return (int)grid.n_keys; }
  void* Context() const { return ctx; }
  Real CellSize() const { return cs; }
  const Domain& GetDomain() const { return domain; }

 private:
  Domain& domain;
  Real cs;
  sph_grid_info grid{};
  std::vector<int> cell_start, cell_end;
  void* ctx = nullptr;
};

// --------------------------------------------------------------- encoder
struct Codifier {  // core/encoders/Encoder.h:11-47
  std::vector<std::vector<int>> neighbors_per_particle;
  void Init(int N) { neighbors_per_particle.resize(N); }
  void ResetCodifier() {
    for (auto& v : neighbors_per_particle) v.clear();
  }
  // receives an ascending list: first index verbatim, then gaps minus one
  void Encode(int pi, std::vector<int>& nb) {
    auto& d = neighbors_per_particle[pi];
    if (nb.empty()) return;
    d.emplace_back(nb[0]);
    for (size_t i = 1; i < nb.size(); ++i) d.emplace_back(nb[i] - nb[i - 1] - 1);
  }
  // inverse of Encode (EncodedListsForceComputation.cpp:23-33)
  std::vector<int> Decode(int pi) const {
    const auto& d = neighbors_per_particle[pi];
    std::vector<int> out(d.size());
    for (size_t i = 0; i < d.size(); ++i) out[i] = i == 0 ? d[0] : out[i - 1] + d[i] + 1;
    return out;
  }
};

// ----------------------------------------------------------- neighbor lists
template <class NeighborSearchType>
class MultipassNeighborLists {
 public:
  explicit MultipassNeighborLists(Real r) : h2(r * r) {}
  // CSR lists, self included (the CPU convention), stencil order
  void AssembleNeighborList(SoAContainer& container,
                            GenericNeighborSearch<NeighborSearchType>& ns) {
    void* ctx = ns.Context();
    if (!ctx) throw sph::Error("FindNeighbors must run first");
    int64_t total = 0;
    sph::check(sph_build_neighbor_lists(ctx, 1, &total), ctx, "sph_build_neighbor_lists");
    size_t n = container.Size();
    std::vector<int64_t> off(n + 1);
    neighbor_list.resize((size_t)total);
    sph::check(sph_get_neighbor_lists(ctx, off.data(), neighbor_list.data()), ctx,
               "sph_get_neighbor_lists");
    offset_per_particle.resize(n + 1);
    for (size_t i = 0; i <= n; ++i) offset_per_particle[i] = (int)off[i];
    total_neighbors = (int)total;
  }
  void ResetNeighborList() {}
  void Resize(int N) { neighbor_list.resize(N); }
  std::vector<int>& GetOffsetPerParticle() { return offset_per_particle; }
  int GetOffsetFromParticle(int i) { return offset_per_particle[i]; }
  std::vector<int>& GetNeighborList() { return neighbor_list; }

 protected:
  Real h2;
  int total_neighbors = 0;
  std::vector<int> offset_per_particle, neighbor_list;
};

using neighbors = std::vector<std::vector<int>>;

template <class NeighborSearchType>
class GenericNeighborLists {
 public:
  explicit GenericNeighborLists(double r) : h2((Real)(r * r)) {}
  // per-particle vectors, self included, stencil order (GenericNeighborLists.tpp:9-81)
  void AssembleNeighborList(SoAContainer& container,
                            GenericNeighborSearch<NeighborSearchType>& ns) {
    MultipassNeighborLists<NeighborSearchType> mp((Real)std::sqrt(h2));
    mp.AssembleNeighborList(container, ns);
    auto& off = mp.GetOffsetPerParticle();
    auto& idx = mp.GetNeighborList();
    size_t n = container.Size();
    neighbor_lists.resize(n);
    for (size_t i = 0; i < n; ++i) neighbor_lists[i].assign(idx.begin() + off[i], idx.begin() + off[i + 1]);
  }
  neighbors& GetNeighborList() { return neighbor_lists; }
  std::vector<int>& GetNeighborsOfParticle(int i) { return neighbor_lists[i]; }
  void ResetNeighborList() {
    for (auto& v : neighbor_lists) v.clear();
  }
  void Resize(int N) { neighbor_lists.resize(N); }
  Codifier& GetCodifier() { return encoder; }

 protected:
  Real h2;
  neighbors neighbor_lists;
  Codifier encoder;
};

// ------------------------------------------------------- force computation
namespace sph {
// density + EOS, then forces, on the GPU; results mirrored into the container.
// GenericForceComputation evaluates over the GPU cell stencil; the list classes
// upload the lists they hold and evaluate those.
inline void GpuDensity(SoAContainer& c, ScalarField& pressure) {
  void* ctx = c.DeviceContext();
  if (!ctx) throw Error("FindNeighbors must run first");
  check(sph_density_pressure(ctx, SPH_COMPUTE_SORTED), ctx, "sph_density_pressure");
  auto& P = c.GetParticles();
  pressure.resize(c.Size());
  check(sph_download_fields(ctx, nullptr, nullptr, P.density.data(), pressure.data()), ctx,
        "sph_download_fields");
}
inline void GpuForces(SoAContainer& c, int compute = SPH_COMPUTE_SORTED) {
  void* ctx = c.DeviceContext();
  if (!ctx) throw Error("FindNeighbors must run first");
  check(sph_forces(ctx, compute), ctx, "sph_forces");
  auto& P = c.GetParticles();
  check(sph_download_fields(ctx, P.fx.data(), P.fy.data(), nullptr, nullptr), ctx,
        "sph_download_fields");
}
// Hands the caller's self-inclusive lists (CSR: offsets[n + 1], indices) to the
// GPU for SPH_COMPUTE_LISTS; they are evaluated as given, like the reference's
// list loops (ListsForceComputation.cpp:3-145, MPListsForceComputation.cpp:3-147).
template <class Off>
inline void UploadLists(SoAContainer& c, const std::vector<Off>& offsets,
                        const std::vector<int>& indices) {
  void* ctx = c.DeviceContext();
  if (!ctx) throw Error("FindNeighbors must run first");
  const size_t n = c.Size();
  if (offsets.size() != n + 1) throw Error("neighbor offsets must hold n + 1 entries");
  std::vector<int64_t> off(offsets.begin(), offsets.end());
  if ((size_t)off[n] > indices.size()) throw Error("neighbor offsets exceed the list");
  check(sph_set_neighbor_lists(ctx, off.data(), indices.data(), 1), ctx,
        "sph_set_neighbor_lists");
}
inline void UploadLists(SoAContainer& c, const std::vector<std::vector<int>>& lists) {
  std::vector<int64_t> off(lists.size() + 1, 0);
  std::vector<int> idx;
  for (size_t i = 0; i < lists.size(); ++i) {
    idx.insert(idx.end(), lists[i].begin(), lists[i].end());
    off[i + 1] = (int64_t)idx.size();
  }
  UploadLists(c, off, idx);
}
inline void GpuListDensity(SoAContainer& c, ScalarField& pressure) {
  void* ctx = c.DeviceContext();
  check(sph_density_pressure(ctx, SPH_COMPUTE_LISTS), ctx, "sph_density_pressure");
  auto& P = c.GetParticles();
  pressure.resize(c.Size());
  check(sph_download_fields(ctx, nullptr, nullptr, P.density.data(), pressure.data()), ctx,
        "sph_download_fields");
}
}  // namespace sph

template <class NeighborSearchType>
class GenericForceComputation {
 public:
  GenericForceComputation(Real radius, GenericNeighborSearch<NeighborSearchType>& ns)
      : h1(radius), neighbor_search(ns) {}
  void Compute(SoAContainer& p) {
    ApproximateDensity(p);
    CalculatePressure(p);
    ComputeForces(p);
  }
  void ApproximateDensity(SoAContainer& p) { sph::GpuDensity(p, pressure_field); }
  void CalculatePressure(SoAContainer&) {}  // computed with the density on the GPU
  void ComputeForces(SoAContainer& p) { sph::GpuForces(p); }
  ScalarField& GetPressure() { return pressure_field; }

 protected:
  Real h1;
  ScalarField pressure_field;
  GenericNeighborSearch<NeighborSearchType>& neighbor_search;
};

// Evaluates the per-particle lists it holds (self included, each particle
// listing itself once), whatever their content: edited or filtered lists give
// the list-based result, as in ListsForceComputation.cpp:3-145.
class ListsForceComputation {
 public:
  ListsForceComputation(Real radius, std::vector<std::vector<int>>& nl) : h1(radius), neighbor_list(nl) {}
  void Compute(SoAContainer& p) {
    ApproximateDensity(p);
    CalculatePressure(p);
    ComputeForces(p);
  }
  void ApproximateDensity(SoAContainer& p) {
    sph::UploadLists(p, neighbor_list);
    sph::GpuListDensity(p, pressure_field);
  }
  void CalculatePressure(SoAContainer&) {}
  void ComputeForces(SoAContainer& p) {
    sph::UploadLists(p, neighbor_list);
    sph::GpuForces(p, SPH_COMPUTE_LISTS);
  }
  ScalarField& GetPressure() { return pressure_field; }

 protected:
  Real h1;
  ScalarField pressure_field;
  std::vector<std::vector<int>>& neighbor_list;
};

// Evaluates the CSR lists it holds (offsets and indices, self included), as
// MPListsForceComputation.cpp:3-147 does.
class MPListsForceComputation {
 public:
  MPListsForceComputation(Real radius, std::vector<int>& nl, std::vector<int>& opp)
      : h1(radius), neighbor_list(nl), offset_per_particle(opp) {}
  void Compute(SoAContainer& p) {
    ApproximateDensity(p);
    CalculatePressure(p);
    ComputeForces(p);
  }
  void ApproximateDensity(SoAContainer& p) {
    sph::UploadLists(p, offset_per_particle, neighbor_list);
    sph::GpuListDensity(p, pressure_field);
  }
  void CalculatePressure(SoAContainer&) {}
  void ComputeForces(SoAContainer& p) {
    sph::UploadLists(p, offset_per_particle, neighbor_list);
    sph::GpuForces(p, SPH_COMPUTE_LISTS);
  }
  ScalarField& GetPressure() { return pressure_field; }

 protected:
  Real h1;
  ScalarField pressure_field;
  std::vector<int>& neighbor_list;
  std::vector<int>& offset_per_particle;
};

// -------------------------------------------------------------- integrator
class Integrator {  // core/integrator/Integrator.h:12-27
 public:
  Integrator(Real delta_t, Domain* domain) : domain(domain), dt(delta_t) {}
  void UpdateVelocities(SoAContainer& c) {
    Ensure(c);
    sph::check(sph_set_dt(c.DeviceContext(), dt), c.DeviceContext(), "sph_set_dt");
    sph::check(sph_update_velocities(c.DeviceContext()), c.DeviceContext(),
               "sph_update_velocities");
  }
  void UpdatePositions(SoAContainer& c) {
    Ensure(c);
    sph::check(sph_set_dt(c.DeviceContext(), dt), c.DeviceContext(), "sph_set_dt");
    sph::check(sph_update_positions(c.DeviceContext()), c.DeviceContext(),
               "sph_update_positions");
    c.PullFromDevice(false);
  }

 private:
  void Ensure(SoAContainer& c) {
    if (!c.DeviceContext()) throw sph::Error("FindNeighbors must run before integrating");
  }
  Domain* domain = nullptr;
  double dt = 0.0;
};

// ------------------------------------------------------------ diagnostics
inline double MeasureGpu(SoAContainer& c, int k) {
  void* ctx = c.DeviceContext();
  if (!ctx) throw sph::Error("no device twin");
  double out[5];
  sph::check(sph_measure(ctx, out), ctx, "sph_measure");
  return out[k];
}
inline Real ComputeKineticEnergy(SoAContainer& c) { return (Real)MeasureGpu(c, 0); }
inline Real ComputeMaxVelocity(SoAContainer& c) { return (Real)MeasureGpu(c, 1); }
inline int ParticlesInDomain(Domain&, SoAContainer& c) { return (int)MeasureGpu(c, 2); }

// The per-step line of the reference driver (app/simulation_configuration.h:152-158),
// from one FP64 measurement on the device.
inline std::string StepHeader(int step, SoAContainer& c, Domain&) {
  void* ctx = c.DeviceContext();
  if (!ctx) throw sph::Error("no device twin");
  double m[5];
  sph::check(sph_measure(ctx, m), ctx, "sph_measure");
  std::ostringstream out;
  out << "Step= " << step << "\t N=" << (int)m[2] << "\t E_k=" << (Real)m[0]
      << "\t |v|_max=" << (Real)m[1] << std::endl;
  return out.str();
}

// Time step from the CFL and force criteria (FP64 maxima on the device).
inline double CflTimeStep(SoAContainer& c, double cfl = 0.25) {
  void* ctx = c.DeviceContext();
  if (!ctx) throw sph::Error("no device twin");
  double out[3];
  sph::check(sph_cfl(ctx, cfl, out), ctx, "sph_cfl");
  return out[0];
}
