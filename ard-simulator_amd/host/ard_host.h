// C++ host mirror of the reference's scene API, running on libard.so.
//
// Same class names, constructors and methods as the reference
// (ARD/simulation.h, ARD/partition.h, ARD/dct_partition.h, ARD/pml_partition.h,
// ARD/boundary.h, ARD/sound_source.h, ARD/gaussian_source.h, ARD/recorder.h),
// so a caller of Simulation::Update keeps its code: the partitions, sources
// and recorders are imported and constructed as before, the Simulation
// constructor builds the same topology (inside the C-ABI), and Update() runs
// one time step of every partition and boundary on the GPU.
//
// Differences, all at the edges of the API:
//  * get_pressure_field() returns a host copy (std::vector) instead of the
//    partition's own array, which lives on the device;
//  * Partition handles of the PML slabs are created by the Simulation
//    constructor and appended to its partition list, as in the reference;
//  * errors throw ard::host::Error instead of asserting or crashing.
#pragma once

#include <cstdint>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ard.h"

namespace ard {
namespace host {

typedef float real_t;  // ARD/types.h:2

struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

class Simulation;
class Recorder;

class Partition {
 public:
  static real_t m_absorption;  // ARD/partition.h:45

  Partition(int xs, int ys, int zs, int w, int h, int d);
  virtual ~Partition() = default;

  // ARD/partition.h:49-57
  virtual bool is_pml() const = 0;
  std::vector<real_t> get_pressure_field();
  std::vector<real_t> get_xy_plane(int z);
  std::vector<real_t> get_yz_plane(int x);
  std::vector<real_t> get_xz_plane(int y);
  real_t get_pressure(int x, int y, int z);
  void set_force(int x, int y, int z, real_t f);
  void Info();

  // ARD/partition.cpp:151-176
  // ARD/partition.cpp:151-176: boxes "x y z w h d" in metres; cell_units reads
  // them as cell indices (seeded scenes whose points are not whole metres)
  static std::vector<std::shared_ptr<Partition>> ImportPartitions(const std::string& path, bool cell_units = false);

  int x_start() const { return x_start_; }
  int y_start() const { return y_start_; }
  int z_start() const { return z_start_; }
  int width() const { return width_; }
  int height() const { return height_; }
  int depth() const { return depth_; }
  int id() const { return id_; }

 protected:
  int x_start_, y_start_, z_start_;
  int width_, height_, depth_;
  int id_ = -1;               // index in the context, set by Simulation
  ard_ctx* ctx_ = nullptr;    // owning context, set by Simulation
  friend class Simulation;
  friend class Recorder;
};

class DctPartition : public Partition {
 public:
  DctPartition(int xs, int ys, int zs, int w, int h, int d) : Partition(xs, ys, zs, w, h, d) {}
  bool is_pml() const override { return false; }
};

class PmlPartition : public Partition {
 public:
  enum PmlType { P_LEFT, P_RIGHT, P_TOP, P_BOTTOM, P_FRONT, P_BACK };  // ARD/pml_partition.h:39-46
  PmlPartition(std::shared_ptr<Partition> neighbor, PmlType type, int xs, int ys, int zs, int w, int h, int d)
      : Partition(xs, ys, zs, w, h, d), neighbor_(neighbor), type_(type) {}
  bool is_pml() const override { return true; }
  PmlType type() const { return type_; }

 private:
  std::shared_ptr<Partition> neighbor_;
  PmlType type_;
};

class SoundSource {
 public:
  // ARD/sound_source.cpp:8-17: numbered in construction order; opens
  // source_<id>.txt in Simulation::OutputDir() (no file if that folder is missing)
  SoundSource(int x, int y, int z);
  virtual ~SoundSource() = default;
  virtual real_t SampleValue(real_t t) = 0;
  // ARD/sound_source.cpp:24-48: also writes each source's source_<id>.txt
  static std::vector<std::shared_ptr<SoundSource>> ImportSources(const std::string& path, bool cell_units = false);
  // ARD/sound_source.cpp:50-57: SampleValue(t) for every step of m_duration, one per line
  void RecordSource();
  int id() const { return id_; }
  int x() const { return x_; }
  int y() const { return y_; }
  int z() const { return z_; }
  // samples for steps 0..n-1 if the source is a fixed waveform; empty for the
  // Gaussian pulse, which the device evaluates itself
  virtual std::vector<float> Waveform() const { return {}; }

 protected:
  int x_, y_, z_;
  int id_;
  std::ofstream source_;
};

class GaussianSource : public SoundSource {
 public:
  GaussianSource(int x, int y, int z) : SoundSource(x, y, z) {}
  real_t SampleValue(real_t t) override;  // ARD/gaussian_source.cpp:16-20
};

// A sampled waveform (band-limited noise drives the gradient checks).
class WaveformSource : public SoundSource {
 public:
  WaveformSource(int x, int y, int z, std::vector<float> w) : SoundSource(x, y, z), w_(std::move(w)) {}
  real_t SampleValue(real_t t) override {
    const int i = int(t);
    return i >= 0 && i < int(w_.size()) ? w_[i] : 0.f;
  }
  std::vector<float> Waveform() const override { return w_; }

 private:
  std::vector<float> w_;
};

class Recorder {
 public:
  // ARD/recorder.cpp:6-17: numbered in construction order; opens out_<id>.txt
  // and response_<id>.txt in Simulation::OutputDir() (no files if that folder
  // is missing, as in the reference)
  Recorder(int x, int y, int z, int total_steps = 1000);
  // ARD/recorder.cpp:26-38
  void FindPartition(const std::vector<std::shared_ptr<Partition>>& partitions);
  // ARD/recorder.cpp:40-57: for time_step < total_steps, appends the 10^3 cube
  // around the recorder to out_<id>.txt and the point sample to
  // response_<id>.txt, in the reference's text format (operator<< on float).
  // Registered with a Simulation, both come from the device traces (no cube
  // rows if the Simulation was built with m_record_cubes = 0).
  void RecordField(int time_step = 0);
  // ARD/recorder.cpp:59-65 (time_step <= total_steps, point sample only)
  void RecordResponse(int time_step = 0);
  void OpenResponse(const std::string& path);  // redirect response_<id>.txt
  void OpenOutput(const std::string& path);    // redirect out_<id>.txt
  int id() const { return id_; }
  static std::vector<std::shared_ptr<Recorder>> ImportRecorders(const std::string& path, bool cell_units = false);
  int x() const { return x_; }
  int y() const { return y_; }
  int z() const { return z_; }

 private:
  int x_, y_, z_, total_steps_;
  int id_;
  std::shared_ptr<Partition> part_;
  Simulation* sim_ = nullptr;
  int device_id_ = -1;  // recorder index in the context when registered
  std::ofstream output_, response_;
  bool sample(int time_step, real_t* point, std::vector<real_t>* cube);
  friend class Simulation;
};

// ARD/boundary.h:9-45: the interface between two partitions, in cell units.
// The Simulation constructor builds one per boundary of its topology, in
// registration (write) order; every boundary's forcing
// (Boundary::ComputeForcingTerms, ARD/boundary.cpp:21-173) runs on the device
// inside Simulation::Update.
class Boundary {
 public:
  enum BoundaryType { X_BOUNDARY, Y_BOUNDARY, Z_BOUNDARY };  // ARD/boundary.h:30-34

  Boundary(BoundaryType type, real_t absorp, std::shared_ptr<Partition> a, std::shared_ptr<Partition> b, int xs,
           int xe, int ys, int ye, int zs, int ze);
  // ARD/boundary.cpp:289-334: the boundary between two partitions that share an
  // x or y face (z extent from a), or nullptr; the geometry of ard_find_boundary
  static std::shared_ptr<Boundary> FindBoundary(std::shared_ptr<Partition> a, std::shared_ptr<Partition> b,
                                                real_t absorp = 1.0f);
  // The forcing of a single boundary is not a separate call here: Update
  // computes every boundary's on the device, in list order. Throws Error.
  void ComputeForcingTerms();
  void Info();  // ARD/boundary.cpp:336-351 format

  BoundaryType type_;
  int id() const { return id_; }
  real_t absorption() const { return absorption_; }
  std::shared_ptr<Partition> a() const { return a_; }
  std::shared_ptr<Partition> b() const { return b_; }
  int x_start() const { return x_start_; }
  int x_end() const { return x_end_; }
  int y_start() const { return y_start_; }
  int y_end() const { return y_end_; }
  int z_start() const { return z_start_; }
  int z_end() const { return z_end_; }

 private:
  int x_start_, x_end_, y_start_, y_end_, z_start_, z_end_;
  int id_;  // numbered in construction order (ARD/boundary.cpp:11-12)
  std::shared_ptr<Partition> a_, b_;
  real_t absorption_;
};

class Simulation {
 public:
  // statics of ARD/simulation.h:41-46, set before constructing partitions
  static real_t m_duration;
  static real_t m_dh;
  static real_t m_dt;
  static real_t m_c0;
  static int m_pml_layers;
  static int m_form;           // ARD_FORM_PLAIN or ARD_FORM_BATCHED
  static int m_recorder_mode;  // ARD_REC_REFERENCE or ARD_REC_LOCAL
  static int m_device;
  static int m_autotune;       // batched form: time the tiling candidates at construction
  static int m_record_cubes;   // registered recorders also sample their 10^3 cube on the device
  static std::string m_output_root;  // "./output/" (ARD/recorder.cpp:13)
  // m_output_root + std::to_string(m_dh) + "_" + std::to_string(m_absorption) + "/"
  // (ARD/recorder.cpp:12-13, ARD/sound_source.cpp:13-14)
  static std::string OutputDir();

  // ARD/simulation.cpp:14-294. `partitions` receives the PML slabs, like the
  // reference's member list. Recorders given here are sampled on the device
  // every step; their RecordField then reads the device trace.
  Simulation(std::vector<std::shared_ptr<Partition>>& partitions,
             std::vector<std::shared_ptr<SoundSource>>& sources,
             std::vector<std::shared_ptr<Recorder>> recorders = {});
  ~Simulation();
  Simulation(const Simulation&) = delete;
  Simulation& operator=(const Simulation&) = delete;

  int Update();  // ARD/simulation.cpp:300-317: returns the step index before the update
  // the lowest partition id whose state holds a NaN or an infinity, or -1
  // (ard_check_finite; no counterpart in the reference)
  int CheckFinite();
  void Info();
  // The visualisation of ARD/simulation.cpp:320-420, drawn on the device when
  // asked for (pixels()) rather than in every Update: look_from_ 0 is the xy
  // plane at the first source's z, 1 the yz plane at its x (ARD/simulation.h:50).
  int look_from_ = 0;
  int size_x();  // ARD/simulation.h:62-73: bounding box of all partitions, PML slabs included
  int size_y();
  int size_z();
  bool ready() const { return ctx_ != nullptr; }
  // the pixel array after drawing the current state (ARD/simulation.h:78-81):
  // size_x * size_y values (xy) or size_y * size_z (yz), RGBA8888 words
  std::vector<uint32_t> pixels();
  ard_ctx* context() { return ctx_; }
  int time_step() const { return time_step_; }
  // the boundaries of the topology in write order (ARD/simulation.h:24-26 m_boundaries)
  const std::vector<std::shared_ptr<Boundary>>& boundaries() const { return boundaries_; }

 private:
  ard_ctx* ctx_ = nullptr;
  std::vector<std::shared_ptr<Partition>> partitions_;
  std::vector<std::shared_ptr<Recorder>> recorders_;
  std::vector<std::shared_ptr<Boundary>> boundaries_;
  int time_step_ = 0;
  // one step's device samples, shared by the recorders' RecordField calls
  int row_t_ = -1;
  std::vector<real_t> row_points_, row_cubes_;
  bool cubes_ = false;
  void fetch_row(int t);
  friend class Recorder;
};

}  // namespace host
}  // namespace ard
