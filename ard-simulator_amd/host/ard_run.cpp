// Command-line driver in the shape of the reference's main loop
// (ARD/main.cpp:47-193) without the SDL window: creates the output folder
// <out_root>/<h>_<absorption>/ (ARD/main.cpp:51-52), imports a scene (the
// import writes source_<id>.txt), builds the Simulation, steps it and, per
// step, lets every recorder write its out_<id>.txt and response_<id>.txt rows.
//
//   ard_run [options] <scene.txt> <sources.txt> [recorders.txt] [steps] [h] [out_root]
//
// out_root defaults to ./output (the folder must exist, as in the reference).
//
// options:
//   --cells       scene, source and recorder files hold cell indices, not metres
//   --rec-local   recorder coordinates are local to their partition (default:
//                 the reference's addressing, global coordinates into the
//                 partition's array)
//   --plain       per-partition form (default: batched)
//   --autotune    batched form: time the tiling candidates before stepping
//   --transform V transform kernels (ard_set_transform_variant: 0 vector FMA,
//                 1 matrix cores, 2 BF16 split, 3 matrix cores even/odd folded)
//   --timing      launch each phase directly and print per-phase kernel times
//   --no-output   step without writing recorder files (no per-step device sync)
//   --no-cubes    recorders write response_<id>.txt only (no 10^3 cube sampling)
//   --c0 V        Simulation::m_c0 in m/s (default 343.5)
//   --duration S  Simulation::m_duration in seconds (default 2; recorders write
//                 int(S / dt) rows, sources int(S / dt) samples)
//   --frame F     after the last step, write the visualisation plane
//                 (Simulation::pixels(), the reference's window contents) to
//                 F as a binary PPM image
//   --look-from V 0: xy plane at the first source's z (default), 1: yz plane at its x
//   --copy-mib N  first run one device-to-device copy of N MiB: a kernel of known
//                 traffic, the scale for hardware byte counters
//   --check-every N  every N steps, check the whole state for NaN or infinity
//                 (ard_check_finite); stop with status 3 at the first non-finite one
//   --info        print every partition and boundary (Partition::Info, Boundary::Info)
//   --ring N      batched form: the modal pass's LDS ring (ard_set_batched_option
//                 ARD_OPT_RING, 0 or 2), for runs that skip the autotuner
//   --pml-shape N batched form: the tiled PML update's tile shape
//                 (ARD_OPT_PML_SHAPE, 0 or 1), likewise
//   --pml-split N batched form: percent of the PML's tile blocks beside
//                 the forward face transforms (ARD_OPT_PML_SPLIT), likewise
#include <hip/hip_runtime_api.h>

#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "ard_host.h"

using namespace ard::host;

int main(int argc, char** argv) {
  bool cells = false, timing = false, output = true;
  long long copy_mib = 0;
  int check_every = 0, ring = -1, pml_shape = -1, pml_split = -1;
  bool info = false;
  const char* frame = nullptr;
  int look_from = 0;
  std::vector<const char*> pos;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--cells")) cells = true;
    else if (!std::strcmp(argv[i], "--rec-local")) Simulation::m_recorder_mode = ARD_REC_LOCAL;
    else if (!std::strcmp(argv[i], "--plain")) Simulation::m_form = ARD_FORM_PLAIN;
    else if (!std::strcmp(argv[i], "--autotune")) Simulation::m_autotune = 1;
    else if (!std::strcmp(argv[i], "--transform") && i + 1 < argc) {
      if (ard_set_transform_variant(std::atoi(argv[++i])) != ARD_OK) {
        std::fprintf(stderr, "%s\n", ard_last_error());
        return 2;
      }
    }
    else if (!std::strcmp(argv[i], "--timing")) timing = true;
    else if (!std::strcmp(argv[i], "--no-output")) output = false;
    else if (!std::strcmp(argv[i], "--no-cubes")) Simulation::m_record_cubes = 0;
    else if (!std::strcmp(argv[i], "--frame") && i + 1 < argc) frame = argv[++i];
    else if (!std::strcmp(argv[i], "--look-from") && i + 1 < argc) look_from = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--c0") && i + 1 < argc) Simulation::m_c0 = float(std::atof(argv[++i]));
    else if (!std::strcmp(argv[i], "--duration") && i + 1 < argc) Simulation::m_duration = float(std::atof(argv[++i]));
    else if (!std::strcmp(argv[i], "--copy-mib") && i + 1 < argc) copy_mib = std::atoll(argv[++i]);
    else if (!std::strcmp(argv[i], "--check-every") && i + 1 < argc) check_every = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--info")) info = true;
    else if (!std::strcmp(argv[i], "--ring") && i + 1 < argc) ring = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--pml-shape") && i + 1 < argc) pml_shape = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--pml-split") && i + 1 < argc) pml_split = std::atoi(argv[++i]);
    else if (argv[i][0] == '-' && argv[i][1] == '-') {
      std::fprintf(stderr, "unknown option %s\n", argv[i]);
      return 2;
    } else {
      pos.push_back(argv[i]);
    }
  }
  if (pos.size() < 2) {
    std::fprintf(stderr,
                 "usage: %s [--cells] [--rec-local] [--plain] [--autotune] [--transform V] [--timing] [--no-output] [--no-cubes] "
                 "[--c0 V] [--duration S] [--frame F] [--look-from V] [--copy-mib N] [--check-every N] [--info] [--ring N] [--pml-shape N] [--pml-split N] scene.txt sources.txt [recorders.txt] [steps] [h] [out_root]\n",
                 argv[0]);
    return 2;
  }
  if (pos.size() > 4) {
    Simulation::m_dh = float(std::atof(pos[4]));
    Simulation::m_dt = Simulation::m_dh < 0.075f ? 0.625e-4f : (Simulation::m_dh < 0.15f ? 1.0f / 8000.0f : 2e-4f);
  }
  if (pos.size() > 5) Simulation::m_output_root = std::string(pos[5]) + "/";
  if (output) mkdir(Simulation::OutputDir().c_str(), 0755);  // the leaf only, like CreateDirectory
  if (copy_mib > 0) {
    const size_t n = size_t(copy_mib) << 20;
    void *a = nullptr, *b = nullptr;
    if (hipSetDevice(Simulation::m_device) != hipSuccess || hipMalloc(&a, n) != hipSuccess ||
        hipMalloc(&b, n) != hipSuccess || hipMemset(a, 1, n) != hipSuccess ||
        hipMemcpy(b, a, n, hipMemcpyDeviceToDevice) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
      std::fprintf(stderr, "calibration copy failed\n");
      return 1;
    }
    hipFree(a);
    hipFree(b);
    std::printf("copied %lld MiB device to device\n", copy_mib);
  }
  try {
    auto t0 = std::chrono::steady_clock::now();
    auto partitions = Partition::ImportPartitions(pos[0], cells);
    auto sources = SoundSource::ImportSources(pos[1], cells);
    std::vector<std::shared_ptr<Recorder>> recorders;
    if (pos.size() > 2) recorders = Recorder::ImportRecorders(pos[2], cells);
    for (auto& r : recorders) r->FindPartition(partitions);
    if (!output) Simulation::m_record_cubes = 0;
    Simulation sim(partitions, sources, recorders);
    const int steps = pos.size() > 3 ? std::atoi(pos[3]) : int(Simulation::m_duration / Simulation::m_dt);
    if (timing && ard_set_kernel_timing(sim.context(), 1) != ARD_OK) throw Error(ard_last_error());
    if (ring >= 0 && ard_set_batched_option(sim.context(), ARD_OPT_RING, ring) != ARD_OK) throw Error(ard_last_error());
    if (pml_shape >= 0 && ard_set_batched_option(sim.context(), ARD_OPT_PML_SHAPE, pml_shape) != ARD_OK)
      throw Error(ard_last_error());
    if (pml_split >= 0 && ard_set_batched_option(sim.context(), ARD_OPT_PML_SPLIT, pml_split) != ARD_OK)
      throw Error(ard_last_error());
    auto t1 = std::chrono::steady_clock::now();
    std::cout << "Initialization finished. (" << std::chrono::duration<double>(t1 - t0).count() << " s)" << std::endl;
    sim.Info();
    if (info) {
      for (auto& p : partitions) p->Info();
      for (auto& b : sim.boundaries()) b->Info();
    }
    for (int s = 0; s < steps; ++s) {
      const int t = sim.Update();
      if (output)
        for (auto& r : recorders) r->RecordField(t);
      if (check_every > 0 && (s + 1) % check_every == 0) {
        const int bad = sim.CheckFinite();
        if (bad >= 0) {
          std::fprintf(stderr, "non-finite state in partition %d after step %d\n", bad, t);
          return 3;
        }
      }
    }
    if (ard_synchronize(sim.context()) != ARD_OK) throw Error(ard_last_error());
    auto t2 = std::chrono::steady_clock::now();
    const double secs = std::chrono::duration<double>(t2 - t1).count();
    std::cout << "Simulation finished. (" << secs << " s, " << steps << " steps)" << std::endl;
    if (frame) {
      sim.look_from_ = look_from;
      const std::vector<uint32_t> px = sim.pixels();
      const int cols = look_from == 1 ? sim.size_y() : sim.size_x();
      const int rows = look_from == 1 ? sim.size_z() : sim.size_y();
      std::FILE* f = std::fopen(frame, "wb");
      if (!f) throw Error(std::string("cannot open ") + frame);
      std::fprintf(f, "P6\n%d %d\n255\n", cols, rows);
      for (uint32_t v : px) {
        const unsigned char rgb[3] = {static_cast<unsigned char>(v >> 16), static_cast<unsigned char>(v >> 8),
                                      static_cast<unsigned char>(v)};
        std::fwrite(rgb, 1, 3, f);
      }
      std::fclose(f);
      std::cout << "Frame written to " << frame << " (" << cols << " x " << rows << ")" << std::endl;
    }
    if (timing) {
      static const char* names[ARD_NUM_PHASES] = {"sources", "pml", "fold", "face_fwd1", "face_fwd2", "modal", "unfold",
                                              "face_inv1", "face_inv2", "interface", "record", "dct_fwd", "dct_inv"};
      float ms[ARD_NUM_PHASES] = {};
      int launches[ARD_NUM_PHASES] = {};
      if (ard_kernel_times(sim.context(), ms, launches, ARD_NUM_PHASES) != ARD_OK) throw Error(ard_last_error());
      for (int p = 0; p < ARD_NUM_PHASES; ++p)
        if (launches[p])
          std::printf("phase %-10s %8.4f ms per launch, %d launches\n", names[p], ms[p], launches[p]);
    }
  } catch (const Error& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
