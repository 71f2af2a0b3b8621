// Timed CPU baseline: the oracle's restatement of the reference's single-rank
// EAM step (half lists, 104-byte AoS, reference phase order), one thread =
// one MPI-rank equivalent.  TEST/BASELINE INFRASTRUCTURE ONLY.
//
// usage: cpu_baseline <setfl> <nx> <ny> <nz> <Fe> <Cu> <Ni> <T> <seed> <steps> [dt]
// prints one JSON line with atom-steps/s and per-phase seconds
// (motion / decide / comm / compute, as simulation.cpp:173-186 splits them).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "oracle.hpp"

int main(int argc, char** argv) {
  if (argc < 11) {
    std::fprintf(stderr, "usage: %s setfl nx ny nz Fe Cu Ni T seed steps [dt]\n", argv[0]);
    return 2;
  }
  const int64_t n[3] = {std::atoll(argv[2]), std::atoll(argv[3]), std::atoll(argv[4])};
  const int ratio[3] = {std::atoi(argv[5]), std::atoi(argv[6]), std::atoi(argv[7])};
  const double T = std::atof(argv[8]);
  const uint32_t seed = static_cast<uint32_t>(std::strtoul(argv[9], nullptr, 10));
  const long steps = std::atol(argv[10]);
  const double dt = argc > 11 ? std::atof(argv[11]) : 0.001;
  const double a = 2.855, f = 5.6 / 2.855;
  oracle::Sim sim;
  std::string err;
  if (!sim.init(n, a, f, argv[1], dt, err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  sim.buildWorld(seed, T, ratio, 1);
  sim.prepareForStart();
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  for (long s = 0; s < steps; s++) sim.step();
  const double secs = std::chrono::duration<double>(clk::now() - t0).count();
  const double atoms = static_cast<double>(sim.globalAtoms());
  std::printf(
      "{\"atoms\": %.0f, \"steps\": %ld, \"seconds\": %.6f, \"atom_steps_per_s\": %.6e, "
      "\"phase_s\": {\"motion\": %.6f, \"decide\": %.6f, \"comm\": %.6f, \"compute\": %.6f}, "
      "\"inter\": %zu, \"T\": %.6f}\n",
      atoms, steps, secs, atoms * steps / secs, sim.t_motion, sim.t_decide, sim.t_comm, sim.t_compute,
      sim.inter_list.size(), sim.temperature());
  return 0;
}
