// Command-line driver in the shape of the reference's main loop
// (ARD/main.cpp:47-193) without the SDL window: imports a scene, builds the
// Simulation, steps it and writes response_<id>.txt for each recorder.
//
//   ard_run <scene.txt> <sources.txt> [recorders.txt] [steps] [h] [out_dir]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "ard_host.h"

using namespace ard::host;

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s scene.txt sources.txt [recorders.txt] [steps] [h] [out_dir]\n", argv[0]);
    return 2;
  }
  if (argc > 5) {
    Simulation::m_dh = float(std::atof(argv[5]));
    Simulation::m_dt = Simulation::m_dh < 0.075f ? 0.625e-4f : (Simulation::m_dh < 0.15f ? 1.0f / 8000.0f : 2e-4f);
  }
  const std::string out_dir = argc > 6 ? argv[6] : ".";
  try {
    auto t0 = std::chrono::steady_clock::now();
    auto partitions = Partition::ImportPartitions(argv[1]);
    auto sources = SoundSource::ImportSources(argv[2]);
    std::vector<std::shared_ptr<Recorder>> recorders;
    if (argc > 3) recorders = Recorder::ImportRecorders(argv[3]);
    for (auto& r : recorders) r->FindPartition(partitions);
    Simulation sim(partitions, sources, recorders);
    for (size_t i = 0; i < recorders.size(); ++i)
      recorders[i]->OpenResponse(out_dir + "/response_" + std::to_string(i) + ".txt");
    const int steps = argc > 4 ? std::atoi(argv[4]) : int(Simulation::m_duration / Simulation::m_dt);
    auto t1 = std::chrono::steady_clock::now();
    std::cout << "Initialization finished. (" << std::chrono::duration<double>(t1 - t0).count() << " s)" << std::endl;
    sim.Info();
    for (int s = 0; s < steps; ++s) {
      const int t = sim.Update();
      for (auto& r : recorders) r->RecordField(t);
    }
    auto t2 = std::chrono::steady_clock::now();
    std::cout << "Simulation finished. (" << std::chrono::duration<double>(t2 - t1).count() << " s, " << steps
              << " steps)" << std::endl;
  } catch (const Error& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
