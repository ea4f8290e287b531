// The reference's CPU driver loop for compute_type=mp-lists, order=morton
// (app/simulation_configuration.h:297-387), written against the drop-in C++ API
// of include/sph/core.hpp, on config/example.txt (96 x 48 box, dx = 0.5,
// 40% x 80% lattice, dt = 0.005). Every stage runs on the GPU. Writes the final
// state and the last neighbor lists to a binary file for the parity test.
#include <cstdio>
#include <algorithm>
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

  // Codifier known answer: first index verbatim, then gaps minus one
  // (core/encoders/Encoder.h:27-44), and its inverse
  Codifier kat;
  kat.Init(1);
  std::vector<int> kat_in{3, 4, 9, 10, 200, 201, 5000};
  const std::vector<int> kat_want{3, 0, 4, 0, 189, 0, 4798};
  kat.Encode(0, kat_in);
  const bool kat_ok = kat.neighbors_per_particle[0] == kat_want && kat.Decode(0) == kat_in;

  auto& P = container.GetParticles();
  const int64_t n = (int64_t)container.Size();
  // final fields of the reference loop, before the edited-list evaluation below
  std::vector<Real> rho_f(P.density), fx_f(P.fx), fy_f(P.fy);
  std::vector<Real> pres_f(computation.GetPressure());

  // Edited lists: drop the last non-self entry of every third particle's row and
  // reverse every fifth row; both list force classes must evaluate these lists,
  // not the cell stencil (ListsForceComputation.cpp:3-145,
  // MPListsForceComputation.cpp:3-147).
  const std::vector<int>& off0 = list.GetOffsetPerParticle();
  const std::vector<int>& idx0 = list.GetNeighborList();
  std::vector<int> eoff(n + 1, 0), eidx;
  std::vector<std::vector<int>> erows(n);
  for (int64_t i = 0; i < n; ++i) {
    std::vector<int> row(idx0.begin() + off0[i], idx0.begin() + off0[i + 1]);
    if (i % 3 == 0)
      for (int k = (int)row.size() - 1; k >= 0; --k)
        if (row[k] != (int)i) {
          row.erase(row.begin() + k);
          break;
        }
    if (i % 5 == 0) std::reverse(row.begin(), row.end());
    erows[i] = row;
    eidx.insert(eidx.end(), row.begin(), row.end());
    eoff[i + 1] = (int)eidx.size();
  }
  MPListsForceComputation edited{H, eidx, eoff};
  edited.ApproximateDensity(container);
  edited.CalculatePressure(container);
  edited.ComputeForces(container);
  std::vector<Real> rho_e(P.density), fx_e(P.fx), fy_e(P.fy);
  ListsForceComputation edited_rows{H, erows};
  edited_rows.Compute(container);
  const bool rows_same = P.density == rho_e && P.fx == fx_e && P.fy == fy_e;
  // a row without its own particle breaks the self-inclusive contract: rejected
  bool rejected = false;
  {
    std::vector<std::vector<int>> bad(erows);
    bad[1].erase(std::find(bad[1].begin(), bad[1].end(), 1));
    ListsForceComputation broken{H, bad};
    try {
      broken.ApproximateDensity(container);
    } catch (const sph::Error& e) {
      rejected = true;
      std::printf("rejected as expected: %s\n", e.what());
    }
  }
  std::ofstream f(out_path, std::ios::binary);
  auto put = [&](const void* p, size_t bytes) { f.write((const char*)p, (std::streamsize)bytes); };
  int64_t total = (int64_t)list.GetNeighborList().size();
  int32_t rt = round_trip ? 1 : 0;
  int32_t flags[3] = {kat_ok ? 1 : 0, rows_same ? 1 : 0, rejected ? 1 : 0};
  put(&n, 8);
  put(&total, 8);
  put(&rt, 4);
  for (auto* v : {&P.rx, &P.ry, &P.vx, &P.vy, &fx_f, &fy_f, &rho_f, &pres_f})
    put(v->data(), (size_t)n * 4);
  put(list.GetOffsetPerParticle().data(), (size_t)(n + 1) * 4);
  put(list.GetNeighborList().data(), (size_t)total * 4);
  put(P.cell_id.data(), (size_t)n * 4);
  // edited lists and what both list classes computed from them
  const int64_t etotal = (int64_t)eidx.size();
  put(flags, sizeof(flags));
  put(&etotal, 8);
  put(eoff.data(), (size_t)(n + 1) * 4);
  put(eidx.data(), (size_t)etotal * 4);
  for (auto* v : {&rho_e, &fx_e, &fy_e}) put(v->data(), (size_t)n * 4);
  std::printf("wrote %s: n=%lld lists=%lld delta round trip=%d codifier KAT=%d\n", out_path,
              (long long)n, (long long)total, rt, flags[0]);
  return round_trip && kat_ok ? 0 : 3;
}
