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
int Simulation::m_record_cubes = 1;
std::string Simulation::m_output_root = "./output/";

std::string Simulation::OutputDir() {
  return m_output_root + std::to_string(m_dh) + "_" + std::to_string(Partition::m_absorption) + "/";
}

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

SoundSource::SoundSource(int x, int y, int z) : x_(x), y_(y), z_(z) {
  static int id_generator = 0;
  id_ = id_generator++;
  source_.open(Simulation::OutputDir() + "source_" + std::to_string(id_) + ".txt");
}

void SoundSource::RecordSource() {
  for (int t = 0; t < Simulation::m_duration / Simulation::m_dt; t++) source_ << SampleValue(real_t(t)) << std::endl;
  source_.close();
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
  for (auto& s : out) s->RecordSource();
  return out;
}

Recorder::Recorder(int x, int y, int z, int total_steps) : x_(x), y_(y), z_(z), total_steps_(total_steps) {
  static int id_generator = 0;
  id_ = id_generator++;
  const std::string dir = Simulation::OutputDir();
  output_.open(dir + "out_" + std::to_string(id_) + ".txt");
  response_.open(dir + "response_" + std::to_string(id_) + ".txt");
}

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
  if (response_.is_open()) response_.close();
  response_.open(path);
  if (!response_) throw Error("cannot open " + path);
}

void Recorder::OpenOutput(const std::string& path) {
  if (output_.is_open()) output_.close();
  output_.open(path);
  if (!output_) throw Error("cannot open " + path);
}

// The point sample and (if asked) the 10^3 cube of `time_step`: from the
// device traces when registered with a Simulation, else read from the
// partition's field after the step, with the reference's addressing (global
// coordinates into the partition's array, ARD/dct_volume.cpp:70-73).
bool Recorder::sample(int time_step, real_t* point, std::vector<real_t>* cube) {
  if (sim_ && device_id_ >= 0) {
    sim_->fetch_row(time_step);
    *point = sim_->row_points_[size_t(device_id_)];
    if (cube) {
      if (!sim_->cubes_) return false;
      cube->assign(sim_->row_cubes_.begin() + size_t(device_id_) * 1000,
                   sim_->row_cubes_.begin() + size_t(device_id_ + 1) * 1000);
    }
    return true;
  }
  if (!part_) throw Error("recorder has no partition (FindPartition found none)");
  const long long W = part_->width_, H = part_->height_;
  long long f0;
  if (Simulation::m_recorder_mode == ARD_REC_LOCAL)
    f0 = ((long long)(z_ - part_->z_start_) * H + (y_ - part_->y_start_)) * W + (x_ - part_->x_start_);
  else
    f0 = ((long long)z_ * H + y_) * W + x_;
  const std::vector<real_t> f = part_->get_pressure_field();
  auto at = [&](long long idx) {
    if (idx < 0 || idx >= (long long)f.size()) throw Error("recorder addresses a cell outside its partition's array");
    return f[size_t(idx)];
  };
  *point = at(f0);
  if (cube) {
    cube->resize(1000);
    for (int i = -5; i < 5; i++)
      for (int j = -5; j < 5; j++)
        for (int k = -5; k < 5; k++) (*cube)[size_t((i + 5) * 100 + (j + 5) * 10 + k + 5)] = at(f0 + (i * H + j) * W + k);
  }
  return true;
}

void Recorder::RecordField(int time_step) {
  if (time_step >= total_steps_) return;
  if (!output_.is_open() && !response_.is_open()) return;
  real_t v;
  std::vector<real_t> cube;
  const bool has_cube = sample(time_step, &v, output_.is_open() ? &cube : nullptr);
  if (output_.is_open() && has_cube) {  // no cube rows when the Simulation samples none (m_record_cubes 0)
    for (real_t c : cube) output_ << c << " ";
    output_ << std::endl;
  }
  if (response_.is_open()) response_ << v << std::endl;
}

void Recorder::RecordResponse(int time_step) {
  if (time_step > total_steps_ || !response_.is_open()) return;
  real_t v;
  sample(time_step, &v, nullptr);
  response_ << v << std::endl;
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
    // boundary handles in write order; those against a PML slab carry the
    // partition absorption (ARD/simulation.cpp:104-260), the others 1
    const int nb = ard_num_boundaries(ctx_);
    for (int k = 0; k < nb; ++k) {
      int b[9];
      check(ard_boundary_info(ctx_, k, b), "ard_boundary_info");
      const auto& pa = partitions_[size_t(b[1])];
      const auto& pb = partitions_[size_t(b[2])];
      const real_t absorp = pa->is_pml() || pb->is_pml() ? Partition::m_absorption : 1.0f;
      boundaries_.push_back(
          std::make_shared<Boundary>(Boundary::BoundaryType(b[0]), absorp, pa, pb, b[3], b[4], b[5], b[6], b[7], b[8]));
    }
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
    cubes_ = m_record_cubes && !recorders_.empty();
    if (cubes_) check(ard_set_recorder_cubes(ctx_, 1), "ard_set_recorder_cubes");
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

void Simulation::fetch_row(int t) {
  if (t == row_t_) return;
  const int nr = ard_num_recorders(ctx_);
  row_points_.resize(size_t(nr));
  check(ard_fetch_recorders(ctx_, row_points_.data(), t, 1), "ard_fetch_recorders");
  if (cubes_) {
    row_cubes_.resize(size_t(nr) * 1000);
    check(ard_fetch_recorder_cubes(ctx_, row_cubes_.data(), t, 1), "ard_fetch_recorder_cubes");
  }
  row_t_ = t;
}

int Simulation::CheckFinite() {
  int bad = -1;
  check(ard_check_finite(ctx_, &bad), "ard_check_finite");
  return bad;
}

int Simulation::Update() {
  const int t = time_step_++;
  check(ard_step(ctx_, 1), "ard_step");
  return t;
}

int Simulation::size_x() {
  int dims[2];
  check(ard_render_view(ctx_, ARD_VIEW_XY, -1, nullptr, dims), "ard_render_view");
  return dims[0];
}
int Simulation::size_y() {
  int dims[2];
  check(ard_render_view(ctx_, ARD_VIEW_XY, -1, nullptr, dims), "ard_render_view");
  return dims[1];
}
int Simulation::size_z() {
  int dims[2];
  check(ard_render_view(ctx_, ARD_VIEW_YZ, -1, nullptr, dims), "ard_render_view");
  return dims[1];
}

std::vector<uint32_t> Simulation::pixels() {
  const int view = look_from_ == 1 ? ARD_VIEW_YZ : ARD_VIEW_XY;
  int dims[2];
  check(ard_render_view(ctx_, view, -1, nullptr, dims), "ard_render_view");
  std::vector<uint32_t> px(size_t(dims[0]) * size_t(dims[1]));
  check(ard_render_view(ctx_, view, -1, px.data(), dims), "ard_render_view");
  return px;
}

Boundary::Boundary(BoundaryType type, real_t absorp, std::shared_ptr<Partition> a, std::shared_ptr<Partition> b,
                   int xs, int xe, int ys, int ye, int zs, int ze)
    : type_(type), x_start_(xs), x_end_(xe), y_start_(ys), y_end_(ye), z_start_(zs), z_end_(ze), a_(std::move(a)),
      b_(std::move(b)), absorption_(absorp) {
  static int id_generator = 0;  // ARD/boundary.cpp:11-12
  id_ = id_generator++;
}

std::shared_ptr<Boundary> Boundary::FindBoundary(std::shared_ptr<Partition> a, std::shared_ptr<Partition> b,
                                                 real_t absorp) {
  const int ba[6] = {a->x_start(), a->y_start(), a->z_start(), a->width(), a->height(), a->depth()};
  const int bb[6] = {b->x_start(), b->y_start(), b->z_start(), b->width(), b->height(), b->depth()};
  int g[9];
  const int r = ard_find_boundary(ba, bb, g);
  check(r, "ard_find_boundary");
  if (r == 0) return nullptr;
  return std::make_shared<Boundary>(BoundaryType(g[0]), absorp, a, b, g[3], g[4], g[5], g[6], g[7], g[8]);
}

void Boundary::ComputeForcingTerms() {
  throw Error("Boundary::ComputeForcingTerms: the forcing of every boundary runs on the device inside "
              "Simulation::Update");
}

void Boundary::Info() {
  const char* t = type_ == X_BOUNDARY ? "X " : (type_ == Y_BOUNDARY ? "Y " : "Z ");
  std::cout << "Boundary " << id_ << ": " << t << a_->id() << "," << b_->id() << " | " << x_start_ << ","
            << y_start_ << "," << z_start_ << "->" << x_end_ << "," << y_end_ << "," << z_end_ << std::endl;
}

void Simulation::Info() {
  std::cout << "Partitions: " << ard_num_partitions(ctx_) << ", boundaries: " << ard_num_boundaries(ctx_)
            << ", recorders: " << ard_num_recorders(ctx_) << std::endl;
}

}  // namespace host
}  // namespace ard
