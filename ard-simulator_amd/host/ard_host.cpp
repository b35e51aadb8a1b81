// C++ host mirror of the reference's scene API over the libard.so C-ABI.
#include "ard_host.h"

#include <cmath>
#include <iostream>

namespace ard {
namespace host {

real_t Partition::m_absorption = 1.0f;  // ARD/main.cpp:29
real_t Simulation::m_duration = 2.0f;   // ARD/main.cpp:30
real_t Simulation::m_dh = 0.1f;         // ARD/main.cpp:38
real_t Simulation::m_dt = 1.0f / 8000.0f;
real_t Simulation::m_c0 = 343.5f;       // ARD/main.cpp:44
int Simulation::m_pml_layers = 5;
int Simulation::m_form = ARD_FORM_BATCHED;
int Simulation::m_recorder_mode = ARD_REC_REFERENCE;
int Simulation::m_device = 0;
int Simulation::m_autotune = 0;

namespace {
void check(int r, const char* what) {
  if (r < 0) throw Error(std::string(what) + ": " + ard_last_error());
}
int m2cell(int v) { return int(float(v) / Simulation::m_dh); }  // (int)(v / m_dh) in FP32
// scene values in metres (the reference's files) or, with cell_units, in cells
int to_cell(int v, bool cell_units) { return cell_units ? v : m2cell(v); }
}  // namespace

Partition::Partition(int xs, int ys, int zs, int w, int h, int d)
    : x_start_(xs), y_start_(ys), z_start_(zs), width_(w), height_(h), depth_(d) {}

std::vector<real_t> Partition::get_pressure_field() {
  if (!ctx_) throw Error("partition not part of a Simulation");
  std::vector<real_t> v(size_t(width_) * height_ * depth_);
  check(ard_copy_pressure_field(ctx_, id_, v.data()), "ard_copy_pressure_field");
  return v;
}

// ARD/partition.cpp:58-89
std::vector<real_t> Partition::get_xy_plane(int z) {
  const std::vector<real_t> f = get_pressure_field();
  return std::vector<real_t>(f.begin() + size_t(z) * width_ * height_, f.begin() + size_t(z + 1) * width_ * height_);
}
std::vector<real_t> Partition::get_yz_plane(int x) {
  const std::vector<real_t> f = get_pressure_field();
  std::vector<real_t> out;
  for (int i = 0; i < depth_; ++i)
    for (int j = 0; j < height_; ++j) out.push_back(f[(size_t(i) * height_ + j) * width_ + x]);
  return out;
}
std::vector<real_t> Partition::get_xz_plane(int y) {
  const std::vector<real_t> f = get_pressure_field();
  std::vector<real_t> out;
  for (int i = 0; i < depth_; ++i)
    for (int j = 0; j < width_; ++j) out.push_back(f[(size_t(i) * height_ + y) * width_ + j]);
  return out;
}

real_t Partition::get_pressure(int x, int y, int z) {
  if (!ctx_) throw Error("partition not part of a Simulation");
  float v = 0.f;
  check(ard_get_pressure(ctx_, id_, x, y, z, &v), "ard_get_pressure");
  return v;
}

void Partition::set_force(int x, int y, int z, real_t f) {
  if (!ctx_) throw Error("partition not part of a Simulation");
  check(ard_set_force(ctx_, id_, x, y, z, f), "ard_set_force");
}

void Partition::Info() {
  std::cout << (is_pml() ? "PML" : "DCT") << " Partition " << id_ << ": " << x_start_ << "," << y_start_ << ","
            << z_start_ << "->" << x_start_ + width_ << "," << y_start_ + height_ << "," << z_start_ + depth_
            << std::endl;
}

std::vector<std::shared_ptr<Partition>> Partition::ImportPartitions(const std::string& path, bool cell_units) {
  std::vector<std::shared_ptr<Partition>> parts;
  std::ifstream file(path);
  if (!file) throw Error("cannot open " + path);
  while (file.good()) {
    int xs, ys, zs, w, h, d;
    file >> xs >> ys >> zs >> w >> h >> d;
    if (file.fail()) break;
    parts.push_back(std::make_shared<DctPartition>(to_cell(xs, cell_units), to_cell(ys, cell_units),
                                                   to_cell(zs, cell_units), to_cell(w, cell_units),
                                                   to_cell(h, cell_units), to_cell(d, cell_units)));
  }
  return parts;
}

real_t GaussianSource::SampleValue(real_t t) {
  const real_t arg = std::pow(float(M_PI) * ((2 * (Simulation::m_c0 * Simulation::m_dt / Simulation::m_dh) * t) / 6 -
                                             2.0f), 2.0f);
  return 1e9f * std::exp(-arg);
}

std::vector<std::shared_ptr<SoundSource>> SoundSource::ImportSources(const std::string& path, bool cell_units) {
  std::vector<std::shared_ptr<SoundSource>> out;
  std::ifstream file(path);
  if (!file) throw Error("cannot open " + path);
  while (file.good()) {
    int x, y, z;
    file >> x >> y >> z;
    if (file.fail()) break;
    out.push_back(std::make_shared<GaussianSource>(to_cell(x, cell_units), to_cell(y, cell_units),
                                                   to_cell(z, cell_units)));
  }
  return out;
}

Recorder::Recorder(int x, int y, int z, int total_steps) : x_(x), y_(y), z_(z), total_steps_(total_steps) {}

void Recorder::FindPartition(const std::vector<std::shared_ptr<Partition>>& partitions) {
  for (auto& p : partitions) {
    if (p->is_pml()) continue;
    if (p->x_start_ < x_ - 5 && p->x_start_ + p->width_ > x_ + 4 && p->y_start_ < y_ - 5 &&
        p->y_start_ + p->height_ > y_ + 4 && p->z_start_ < z_ - 5 && p->z_start_ + p->depth_ > z_ + 4) {
      part_ = p;
      return;
    }
  }
}

void Recorder::OpenResponse(const std::string& path) {
  response_.open(path);
  if (!response_) throw Error("cannot open " + path);
}

void Recorder::RecordField(int time_step) {
  if (time_step >= total_steps_) return;
  float v;
  if (sim_ && device_id_ >= 0) {
    const int nr = ard_num_recorders(sim_->ctx_);
    std::vector<float> row(static_cast<size_t>(nr));
    check(ard_fetch_recorders(sim_->ctx_, row.data(), time_step, 1), "ard_fetch_recorders");
    v = row[size_t(device_id_)];
  } else {
    if (!part_) throw Error("recorder has no partition (FindPartition found none)");
    // the reference's addressing: global coordinates into the partition's array
    const long long flat = (long long)z_ * part_->height_ * part_->width_ + (long long)y_ * part_->width_ + x_;
    const int lx = int(flat % part_->width_), ly = int(flat / part_->width_ % part_->height_),
              lz = int(flat / ((long long)part_->width_ * part_->height_));
    v = part_->get_pressure(lx, ly, lz);
  }
  if (response_.is_open()) response_ << v << std::endl;
}

std::vector<std::shared_ptr<Recorder>> Recorder::ImportRecorders(const std::string& path, bool cell_units) {
  std::vector<std::shared_ptr<Recorder>> out;
  std::ifstream file(path);
  if (!file) throw Error("cannot open " + path);
  while (file.good()) {
    int x, y, z;
    file >> x >> y >> z;
    if (file.fail()) break;
    out.push_back(std::make_shared<Recorder>(to_cell(x, cell_units), to_cell(y, cell_units), to_cell(z, cell_units),
                                             int(Simulation::m_duration / Simulation::m_dt)));
  }
  return out;
}

Simulation::Simulation(std::vector<std::shared_ptr<Partition>>& partitions,
                       std::vector<std::shared_ptr<SoundSource>>& sources,
                       std::vector<std::shared_ptr<Recorder>> recorders)
    : recorders_(std::move(recorders)) {
  ard_params p{};
  p.h = m_dh; p.dt = m_dt; p.c0 = m_c0; p.pml_layers = m_pml_layers; p.absorption = Partition::m_absorption;
  p.form = m_form; p.recorder_mode = m_recorder_mode; p.autotune = m_autotune;
  check(ard_ctx_create(m_device, &p, &ctx_), "ard_ctx_create");
  try {
    for (auto& part : partitions) {
      if (part->is_pml()) throw Error("PML partitions are created by the Simulation constructor");
      part->id_ = ard_add_dct_partition(ctx_, part->x_start_, part->y_start_, part->z_start_, part->width_,
                                        part->height_, part->depth_);
      check(part->id_, "ard_add_dct_partition");
      part->ctx_ = ctx_;
    }
    check(ard_build_topology(ctx_), "ard_build_topology");
    const int np = ard_num_partitions(ctx_);
    for (int i = int(partitions.size()); i < np; ++i) {
      int info[9];
      check(ard_partition_info(ctx_, i, info), "ard_partition_info");
      auto pml = std::make_shared<PmlPartition>(partitions[size_t(info[2])], PmlPartition::PmlType(info[1]), info[3],
                                                info[4], info[5], info[6], info[7], info[8]);
      pml->id_ = i;
      pml->ctx_ = ctx_;
      partitions.push_back(pml);
    }
    partitions_ = partitions;
    const int steps = int(m_duration / m_dt);
    for (auto& s : sources) {
      std::vector<float> w = s->Waveform();
      if (w.empty() && !dynamic_cast<GaussianSource*>(s.get())) {
        w.resize(size_t(steps));
        for (int t = 0; t < steps; ++t) w[size_t(t)] = s->SampleValue(real_t(t));
      }
      check(ard_add_source(ctx_, s->x(), s->y(), s->z(), w.empty() ? nullptr : w.data(), int(w.size())),
            "ard_add_source");
    }
    for (auto& r : recorders_) {
      r->device_id_ = ard_add_recorder(ctx_, r->x_, r->y_, r->z_);
      check(r->device_id_, "ard_add_recorder");
      r->sim_ = this;
    }
    check(ard_finalize(ctx_), "ard_finalize");
  } catch (...) {
    ard_ctx_destroy(ctx_);
    ctx_ = nullptr;
    throw;
  }
}

Simulation::~Simulation() {
  for (auto& r : recorders_) r->sim_ = nullptr;
  for (auto& p : partitions_) p->ctx_ = nullptr;
  if (ctx_) ard_ctx_destroy(ctx_);
}

int Simulation::Update() {
  const int t = time_step_++;
  check(ard_step(ctx_, 1), "ard_step");
  return t;
}

void Simulation::Info() {
  std::cout << "Partitions: " << ard_num_partitions(ctx_) << ", boundaries: " << ard_num_boundaries(ctx_)
            << ", recorders: " << ard_num_recorders(ctx_) << std::endl;
}

}  // namespace host
}  // namespace ard
