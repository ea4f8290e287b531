// The reference's CPU driver loop for compute_type=mp-lists, order=morton
// (app/simulation_configuration.h:297-387), written against the drop-in C++ API
// of include/sph/core.hpp, on config/example.txt (96 x 48 box, dx = 0.5,
// 40% x 80% lattice, dt = 0.005). Every stage runs on the GPU. Writes the final
// state and the last neighbor lists to a binary file for the parity test.
#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "sph/core.hpp"

int main(int argc, char** argv) {
  const char* out_path = argc > 1 ? argv[1] : "core_api_example.bin";
  const int steps = argc > 2 ? std::atoi(argv[2]) : 5;
  const Real dt = 0.005f;

  Domain domain;
  domain.l_corner = {0, 0, 0};
  domain.h_corner = {96, 48, 0};
  SoAContainer container;
  LatticeCreator lattice;
  lattice.xVol = 0.4; lattice.yVol = 0.8; lattice.dX = 0.5; lattice.dY = 0.5;
  lattice.CreateLattice(domain, container);
  container.dt_hint = dt;

  GenericNeighborSearch<MortonOrder> gns{domain, H};
  MultipassNeighborLists<MortonOrder> list{H};
  MPListsForceComputation computation{H, list.GetNeighborList(), list.GetOffsetPerParticle()};

  // pre-loop (PreLoopComputationWithMPLists)
  gns.FindNeighbors(container);
  list.AssembleNeighborList(container, gns);
  computation.ApproximateDensity(container);
  computation.CalculatePressure(container);
  computation.ComputeForces(container);
  std::printf("Step= 0\t N=%d\t E_k=%g\t |v|_max=%g\n", ParticlesInDomain(domain, container),
              ComputeKineticEnergy(container), ComputeMaxVelocity(container));

  Integrator integrator{dt, &domain};
  for (int i = 1; i < steps; ++i) {  // MainLoopComputationWithMPLists
    integrator.UpdateVelocities(container);
    integrator.UpdatePositions(container);
    gns.FindNeighbors(container);
    list.AssembleNeighborList(container, gns);
    computation.ApproximateDensity(container);
    computation.CalculatePressure(container);
    computation.ComputeForces(container);
    std::printf("Step= %d\t N=%d\t E_k=%g\t |v|_max=%g\n", i,
                ParticlesInDomain(domain, container), ComputeKineticEnergy(container),
                ComputeMaxVelocity(container));
  }
  // the same line from the one-measurement helper, and the CFL time step
  std::fputs(StepHeader(steps - 1, container, domain).c_str(), stdout);
  std::printf("cfl dt=%.9g\n", CflTimeStep(container, 0.25));

  // delta-encode the first particle's (sorted) list and decode it back
  GenericNeighborLists<MortonOrder> vlists{H};
  vlists.AssembleNeighborList(container, gns);
  Codifier& cod = vlists.GetCodifier();
  cod.Init((int)container.Size());
  std::vector<int> l0 = vlists.GetNeighborsOfParticle(0);
  std::sort(l0.begin(), l0.end());
  cod.Encode(0, l0);
  bool round_trip = cod.Decode(0) == l0;

  auto& P = container.GetParticles();
  const int64_t n = (int64_t)container.Size();
  std::ofstream f(out_path, std::ios::binary);
  auto put = [&](const void* p, size_t bytes) { f.write((const char*)p, (std::streamsize)bytes); };
  int64_t total = (int64_t)list.GetNeighborList().size();
  int32_t rt = round_trip ? 1 : 0;
  put(&n, 8);
  put(&total, 8);
  put(&rt, 4);
  for (auto* v : {&P.rx, &P.ry, &P.vx, &P.vy, &P.fx, &P.fy, &P.density})
    put(v->data(), (size_t)n * 4);
  put(computation.GetPressure().data(), (size_t)n * 4);
  put(list.GetOffsetPerParticle().data(), (size_t)(n + 1) * 4);
  put(list.GetNeighborList().data(), (size_t)total * 4);
  put(P.cell_id.data(), (size_t)n * 4);
  std::printf("wrote %s: n=%lld lists=%lld delta round trip=%d\n", out_path, (long long)n,
              (long long)total, rt);
  return round_trip ? 0 : 3;
}
