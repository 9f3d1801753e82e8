// fdtd_app.cpp -- FDTD::App facade over the C ABI. Reference behaviour cited per function.
#include "fdtd_app.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "pfdtd.h"

namespace {

void log_line(const char *what, const char *msg) {
  std::fprintf(stderr, "[FDTD::App] %s: %s\n", what, msg);
}

// Defaults of App::setupDefaultCallbacks (src/App.cpp:55-86): never interrupt, print every
// 100th step.
bool default_interrupt() { return false; }
void default_progress(int step, int max_step, float t_per_step) {
  std::printf("Step %d/%d, time per step %f, estimated time left %f s \n", step, max_step,
              t_per_step, t_per_step * (float)(max_step - step));
}

int interrupt_tramp(void *user) {
  FDTD::App *app = static_cast<FDTD::App *>(user);
  return app->m_interrupt && app->m_interrupt() ? 1 : 0;
}
void progress_tramp(int step, int n, float sec, void *user) {
  FDTD::App *app = static_cast<FDTD::App *>(user);
  if (app->m_progress) app->m_progress(step, n, sec);
}

}  // namespace

// loggerInit, src/logger.cpp: start an empty solver_log.txt in the working directory
void loggerInit() {
  std::wofstream logfile("solver_log.txt", std::fstream::out | std::fstream::trunc);
}

// ---------------------------------------------------------------------------------------
// Position / SimulationParameters
// ---------------------------------------------------------------------------------------

// Position::getElementIdx, src/base/SrcRec.cpp:26-34
nv::Vec3i Position::getElementIdx(unsigned fs, float c, float lambda) const {
  const float dx = c / ((float)fs * lambda);
  return nv::Vec3i((int)std::floor(p_.x / dx + 0.5f), (int)std::floor(p_.y / dx + 0.5f),
                   (int)std::floor(p_.z / dx + 0.5f));
}

// SimulationParameters::readGridIr + FileReader::readFloat (src/io/FileReader.cpp:106-127)
void SimulationParameters::readGridIr(const std::string &path) {
  grid_ir_.clear();
  std::ifstream in(path);
  if (!in.good()) {
    log_line("readGridIr", ("cannot open " + path).c_str());
    return;
  }
  float v;
  while (in >> v) grid_ir_.push_back(v);
}

// setUpdateType, SimulationParameters.cpp:123-137: lambda = sqrt(1/3) for every type
void SimulationParameters::setUpdateType(UpdateType t) {
  update_type_ = t;
  lambda_ = std::sqrt(1.0 / 3.0);
}

// getDx, SimulationParameters.cpp:362-364
float SimulationParameters::getDx() const {
  return (float)((double)c_ / ((double)spatial_fs_ * lambda_));
}

void SimulationParameters::removeSource(unsigned int i) {
  if (i >= sources_.size()) throw std::out_of_range("SimulationParameters::removeSource: idx out of range");
  sources_.erase(sources_.begin() + i);
}

void SimulationParameters::removeReceiver(unsigned int i) {
  if (i >= receivers_.size()) throw std::out_of_range("SimulationParameters::removeReceiver: idx out of range");
  receivers_.erase(receivers_.begin() + i);
}

// addInputData(float*, n), SimulationParameters.cpp:96-113
void SimulationParameters::addInputData(const float *data, unsigned int n) {
  std::vector<float> v(n, 0.f);
  if (data)
    for (unsigned int i = 0; i < n; ++i) v[i] = data[i];
  input_f_.push_back(v);
}

// getInputDataSample(+Double), SimulationParameters.cpp:161-189: 0 past the end
float SimulationParameters::getInputDataSample(unsigned int idx, unsigned int sample) const {
  const std::vector<float> &v = input_f_.at(idx);
  return sample < v.size() ? v[sample] : 0.f;
}
double SimulationParameters::getInputDataSampleDouble(unsigned int idx, unsigned int sample) const {
  const std::vector<double> &v = input_d_.at(idx);
  return sample < v.size() ? v[sample] : 0.0;
}

float SimulationParameters::getGridIrDataSample(unsigned int sample) const {
  return sample < grid_ir_.size() ? grid_ir_[sample] : 0.f;
}

// getRegularSourceSample, SimulationParameters.cpp:262-298
float SimulationParameters::regularSample(unsigned int src, unsigned int step) const {
  const Source s = sources_.at(src);
  float sample = 0.f;
  switch (s.getInputType()) {
  case IMPULSE:
    sample += step == 1 ? 1.f : 0.f;
    break;
  case GAUSSIAN: {
    const float t0 = 40, width = 4;
    const float e = ((float)step - t0) / width;
    sample += std::exp(-0.5f * (e * e));
    break;
  }
  case DATA:
    sample += getInputDataSample(s.getInputDataIdx(), step);
    break;
  case SINE: {
    const float freq = 120;
    const float t = (float)step / (float)spatial_fs_;
    sample += std::sin(2.f * (float)3.14159265358979323846 * freq * t);
    break;
  }
  }
  return sample;
}

// getRegularSourceSampleDouble, SimulationParameters.cpp:312-348
double SimulationParameters::regularSampleDouble(unsigned int src, unsigned int step) const {
  const Source s = sources_.at(src);
  double sample = 0.0;
  switch (s.getInputType()) {
  case IMPULSE:
    sample += step == 1 ? 1.0 : 0.0;
    break;
  case GAUSSIAN: {
    const double t0 = 40, width = 4;
    const double e = ((float)((double)step - t0) / width);
    sample += std::exp(-0.5f * (e * e));
    break;
  }
  case DATA:
    sample += getInputDataSampleDouble(s.getInputDataIdx(), step);
    break;
  case SINE: {
    const double freq = 120;
    const double t = (double)step / (double)spatial_fs_;
    sample += std::sin(2.f * (double)3.14159265358979323846 * freq * t);
    break;
  }
  }
  return sample;
}

// getSourceSample(+Double), SimulationParameters.cpp:139-159, with the transparent source
// s(n) - sum_{i<n} h(n-i) s(i) of :300-310 and :350-360
float SimulationParameters::getSourceSample(unsigned int src, unsigned int step) const {
  if (sources_.at(src).getSourceType() != SRC_TRANSPARENT) return 0.f + regularSample(src, step);
  float acc = 0.f;
  for (unsigned int i = 0; i < step; ++i) acc += getGridIrDataSample(step - i) * regularSample(src, i);
  return 0.f + (regularSample(src, step) - acc);
}

double SimulationParameters::getSourceSampleDouble(unsigned int src, unsigned int step) const {
  if (sources_.at(src).getSourceType() != SRC_TRANSPARENT) return 0.0 + regularSampleDouble(src, step);
  double acc = 0.0;
  for (unsigned int i = 0; i < step; ++i)
    acc += getGridIrDataSample(step - i) * regularSampleDouble(src, i);
  return 0.0 + (regularSampleDouble(src, step) - acc);
}

// getSource/ReceiverElementCoordinates, SimulationParameters.cpp:200-224
nv::Vec3i SimulationParameters::getSourceElementCoordinates(unsigned int i) const {
  nv::Vec3i r = getSource(i).getElementIdx(spatial_fs_, c_, (float)lambda_);
  if (add_padding_to_element_idx_) {
    ++r.x;
    ++r.y;
    ++r.z;
  }
  return r;
}

nv::Vec3i SimulationParameters::getReceiverElementCoordinates(unsigned int i) const {
  nv::Vec3i r = getReceiver(i).getElementIdx(spatial_fs_, c_, (float)lambda_);
  if (add_padding_to_element_idx_) {
    ++r.x;
    ++r.y;
    ++r.z;
  }
  return r;
}

std::vector<float> SimulationParameters::getParameterVector() const {
  return {(float)lambda_, (float)(lambda_ * lambda_), 1.f / 3.f, (float)octave_};
}

std::vector<double> SimulationParameters::getParameterVectorDouble() const {
  return {lambda_, lambda_ * lambda_, (double)1 / (double)3, (double)octave_};
}

// ---------------------------------------------------------------------------------------
// MaterialHandler (src/base/MaterialHandler.cpp)
// ---------------------------------------------------------------------------------------

// addMaterials, MaterialHandler.cpp:27-45
void MaterialHandler::addMaterials(const float *coefs, unsigned int n_surfaces,
                                   unsigned int n_coefs) {
  std::vector<float> tmp(kCoefNum, 0.f);
  for (unsigned int i = 0; i < n_surfaces; ++i) {
    for (unsigned int j = 0; j < n_coefs && j < kCoefNum; ++j) tmp[j] = coefs[(size_t)i * n_coefs + j];
    addSurfaceMaterial(tmp);
  }
}

// addSurfaceMaterial + checkIfUnique, MaterialHandler.cpp:47-66, :186-204
unsigned int MaterialHandler::addSurfaceMaterial(std::vector<float> coefs) {
  coefs.resize(kCoefNum, 0.f);
  unsigned int idx = 0;
  for (; idx < unique_.size(); ++idx)
    if (std::equal(coefs.begin(), coefs.end(), unique_[idx].begin())) break;
  if (idx == unique_.size()) unique_.push_back(coefs);
  indices_.push_back((unsigned char)idx);
  return idx;
}

// setGlobalMaterial, MaterialHandler.cpp:206-216
void MaterialHandler::setGlobalMaterial(unsigned int n_surfaces, float coef) {
  std::vector<float> m(kCoefNum, coef);
  for (unsigned int i = 0; i < n_surfaces; ++i) addSurfaceMaterial(m);
}

// getUniqueCoefAt / getSurfaceCoefAt, MaterialHandler.cpp:68-89
float MaterialHandler::getUniqueCoefAt(unsigned int m, unsigned int c) const {
  if (c >= kCoefNum) throw std::out_of_range("Coef idx out of range");
  return unique_.at(m).at(c);
}
float MaterialHandler::getSurfaceCoefAt(unsigned int s, unsigned int c) const {
  if (c >= kCoefNum) throw std::out_of_range("Coef idx out of range");
  return getUniqueCoefAt(indices_.at(s), c);
}

// getMaterialCoefficientPtr, MaterialHandler.cpp:100-130 (reflectances converted to
// admittances when flagged)
std::vector<float> MaterialHandler::getMaterialCoefficients() const {
  std::vector<float> t;
  t.reserve(unique_.size() * kCoefNum);
  for (const auto &m : unique_)
    for (unsigned int j = 0; j < kCoefNum; ++j) t.push_back(admittance_ ? m[j] : reflection2Admitance(m[j]));
  return t;
}

// getMeanAbsorption, MaterialHandler.cpp:166-174
float MaterialHandler::getMeanAbsorption(unsigned int octave) const {
  float acc = 0.f;
  for (unsigned int i = 0; i < indices_.size(); ++i) {
    const float r = admitance2Reflection(getSurfaceCoefAt(i, octave));
    acc += 1 - r * r;
  }
  return indices_.empty() ? 0.f : acc / (float)indices_.size();
}

// ---------------------------------------------------------------------------------------
// GeometryHandler
// ---------------------------------------------------------------------------------------
void GeometryHandler::initialize(const unsigned int *indices, const float *vertices,
                                 unsigned int n_idx, unsigned int n_vert) {
  indices_.assign(indices, indices + n_idx);
  vertices_.assign(vertices, vertices + n_vert);
}

void GeometryHandler::setLayerIndices(const std::vector<int> &idx, const std::string &name) {
  layers_.push_back({name, idx});
}

// ---------------------------------------------------------------------------------------
// App
// ---------------------------------------------------------------------------------------
namespace FDTD {

App::App() { setupDefaultCallbacks(); }

App::~App() {
  if (ctx_) pfdtd_ctx_destroy(ctx_);
}

void App::setupDefaultCallbacks() {
  m_interrupt = default_interrupt;
  m_progress = default_progress;
}

void App::check(int rc, const char *what) {
  if (rc != PFDTD_OK) {
    log_line(what, pfdtd_last_error());
    throw(-1);
  }
}

// initializeDevices, src/App.cpp:125-131: one context on one device, the first visible one
// (HIP_VISIBLE_DEVICES selects it; the reference queries and resets every device and picks the
// highest-GFLOPS one). Only that device is touched, and it is not reset.
void App::initializeDevices() {
  int n = 0;
  check(pfdtd_device_count(&n), "initializeDevices");
  if (n < 1) {
    log_line("initializeDevices", "no device");
    throw(-1);
  }
  device_ = 0;
  if (ctx_) {
    pfdtd_ctx_destroy(ctx_);
    ctx_ = nullptr;
    has_domain_ = false;
  }
  check(pfdtd_ctx_create(device_, &ctx_), "initializeDevices");
}

void App::ensureContext() {
  if (!ctx_) initializeDevices();
}

void App::initializeGeometry(const unsigned int *indices, const float *vertices,
                             unsigned int number_of_indices, unsigned int number_of_vertices) {
  m_geometry.initialize(indices, vertices, number_of_indices, number_of_vertices);
  geometry_domain_ = true;
  has_domain_ = false;
  labels_.clear();
}

// FileReader::readVTK, src/io/FileReader.cpp:41-89: among the first 20 lines, "POINTS n ..."
// is followed by 3n coordinates in inches, converted to metres and rounded to
// truncate_vertices_to = 0.01 m (ROUND(v*0.0254/t)*t in float, geomMath.h:126-128), and
// "POLYGONS n size" by n rows "3 i1 i2 i3".
void App::initializeGeometryFromFile(const std::string &path) {
  std::ifstream in(path);
  if (!in.good()) {
    log_line("initializeGeometryFromFile", ("invalid file: " + path).c_str());
    throw(-1);
  }
  const float trunc = 0.01f;
  std::vector<float> vertices;
  std::vector<unsigned int> indices;
  std::string line;
  for (int count = 0; count < 20 && std::getline(in, line); ++count) {
    std::istringstream ss(line);
    std::string key;
    ss >> key;
    if (key == "POINTS") {
      unsigned int n = 0;
      ss >> n;
      vertices.assign((size_t)n * 3, 0.f);
      for (size_t i = 0; i < vertices.size() && in; ++i) {
        float v = 0.f;
        in >> v;
        v = std::floor((v * 0.0254f) / trunc + 0.5f);
        vertices[i] = v * trunc;
      }
    } else if (key == "POLYGONS") {
      unsigned int n = 0;
      ss >> n;
      indices.assign((size_t)n * 3, 0u);
      for (size_t i = 0; i < indices.size() && in; i += 3) {
        unsigned int k = 0;
        in >> k >> indices[i] >> indices[i + 1] >> indices[i + 2];
      }
    }
  }
  if (!in && !in.eof()) {
    log_line("initializeGeometryFromFile", ("malformed file: " + path).c_str());
    throw(-1);
  }
  initializeGeometry(indices.data(), vertices.data(), (unsigned int)indices.size(),
                     (unsigned int)vertices.size());
}

void App::initializeVoxels(const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y,
                           int64_t Z) {
  ensureContext();
  // position bytes are in the encoding of the current update type (toBilbao / toKowalczyk)
  check(pfdtd_set_update_type(ctx_, (int)m_parameters.getUpdateType()), "initializeVoxels");
  check(pfdtd_set_domain(ctx_, pos, mat, X, Y, Z), "initializeVoxels");
  labels_.clear();
  geometry_domain_ = false;
  domain_type_ = (int)m_parameters.getUpdateType();
  uint64_t air = 0, bnd = 0;
  check(pfdtd_get_counts(ctx_, &air, &bnd), "initializeVoxels");
  m_mesh.set(X, Y, Z, air, bnd);
  has_domain_ = true;
}

void App::initializeVoxelLabels(const uint8_t *labels, int64_t X, int64_t Y, int64_t Z) {
  ensureContext();
  check(pfdtd_set_update_type(ctx_, (int)m_parameters.getUpdateType()), "initializeVoxelLabels");
  check(pfdtd_set_domain_labels(ctx_, labels, X, Y, Z), "initializeVoxelLabels");
  // kept so that a later setUpdateType can reclassify (the reference classifies for the
  // update type when the mesh is initialised, cudaMesh.cu:66-74)
  labels_.assign(labels, labels + (size_t)(X * Y * Z));
  geometry_domain_ = false;
  labels_dims_[0] = X;
  labels_dims_[1] = Y;
  labels_dims_[2] = Z;
  domain_type_ = (int)m_parameters.getUpdateType();
  uint64_t air = 0, bnd = 0;
  check(pfdtd_get_counts(ctx_, &air, &bnd), "initializeVoxelLabels");
  m_mesh.set(X, Y, Z, air, bnd);
  has_domain_ = true;
}

// The device-side half of App::initializeMesh (src/App.cpp:147-234): materials and
// parameters. The voxel domain is already resident.
void App::initializeMesh() {
  ensureContext();
  const int type = (int)m_parameters.getUpdateType();
  if (geometry_domain_) {
    // voxelizeGeometry + setupMesh (App.cpp:178-211): voxelize at dx, pad to (32,4,1)
    const unsigned int n_tri = m_geometry.getNumberOfTriangles();
    if (n_tri == 0 || m_materials.getNumberOfSurfaces() < n_tri) {
      log_line("initializeMesh", "geometry needs a material index per triangle "
                                 "(addSurfaceMaterials / setUniformMaterial)");
      throw(-1);
    }
    std::vector<uint8_t> tmat(n_tri);
    for (unsigned int t = 0; t < n_tri; ++t) tmat[t] = (uint8_t)m_materials.getMaterialIdxAt(t);
    const int32_t pad[3] = {32, 4, 1};
    int64_t dims[3] = {0, 0, 0};
    check(pfdtd_set_update_type(ctx_, type), "initializeMesh");
    check(pfdtd_voxelize(ctx_, m_geometry.vertices().data(),
                         (int64_t)m_geometry.getNumberOfVertices(), m_geometry.indices().data(),
                         (int64_t)n_tri, tmat.data(), m_parameters.getDx(), pad, dims, nullptr),
          "initializeMesh");
    uint64_t air = 0, bnd = 0;
    check(pfdtd_get_counts(ctx_, &air, &bnd), "initializeMesh");
    m_mesh.set(dims[0], dims[1], dims[2], air, bnd);
    domain_type_ = type;
    has_domain_ = true;
  }
  if (!has_domain_) {
    log_line("initializeMesh",
             "no domain: call initializeGeometry/initializeGeometryFromFile or "
             "initializeVoxels/initializeVoxelLabels");
    throw(-1);
  }
  if ((type == SRL) != (domain_type_ == SRL)) {
    // the boundary formulation changed since the domain was classified
    if (labels_.empty()) {
      log_line("initializeMesh",
               "update type changed after initializeVoxels; pass position bytes of the new "
               "type's encoding");
      throw(-1);
    }
    check(pfdtd_set_update_type(ctx_, type), "initializeMesh");
    check(pfdtd_set_domain_labels(ctx_, labels_.data(), labels_dims_[0], labels_dims_[1],
                                  labels_dims_[2]),
          "initializeMesh");
    uint64_t air = 0, bnd = 0;
    check(pfdtd_get_counts(ctx_, &air, &bnd), "initializeMesh");
    m_mesh.set(labels_dims_[0], labels_dims_[1], labels_dims_[2], air, bnd);
    domain_type_ = type;
  } else {
    check(pfdtd_set_update_type(ctx_, type), "initializeMesh");
  }
  const std::vector<float> table = m_materials.getMaterialCoefficients();
  if (table.empty()) {
    log_line("initializeMesh", "no materials assigned");
    throw(-1);
  }
  check(pfdtd_set_materials(ctx_, table.data(), (int)m_materials.getNumberOfUniqueMaterials(),
                            (int)MaterialHandler::kCoefNum, m_parameters.getOctave(),
                            m_parameters.getLambda(), quirk_flags_),
        "initializeMesh");
}

// Sources (signals precomputed for every step, FP32 and FP64 as getSourceSample(+Double)
// give them) and receivers, at the padded element coordinates.
void App::uploadSourcesAndReceivers(bool single_precision_only) {
  const unsigned int ns = m_parameters.getNumSources();
  const unsigned int nr = m_parameters.getNumReceivers();
  const unsigned int steps = m_parameters.getNumSteps();
  std::vector<int32_t> sxyz(3 * (size_t)ns), kind(ns), rxyz(3 * (size_t)nr);
  std::vector<float> sf((size_t)ns * steps);
  std::vector<double> sd(single_precision_only ? 0 : (size_t)ns * steps);
  for (unsigned int s = 0; s < ns; ++s) {
    const nv::Vec3i q = m_parameters.getSourceElementCoordinates(s);
    sxyz[3 * s] = q.x;
    sxyz[3 * s + 1] = q.y;
    sxyz[3 * s + 2] = q.z;
    kind[s] = (int32_t)m_parameters.getSource(s).getSourceType();
    // regular samples here; transparent sources are filtered on the device below (the same
    // values getSourceSample(+Double) returns, without the O(steps^2) host loop)
    for (unsigned int n = 0; n < steps; ++n) {
      sf[(size_t)s * steps + n] = 0.f + m_parameters.getRegularSourceSample(s, n);
      if (!single_precision_only)
        sd[(size_t)s * steps + n] = 0.0 + m_parameters.getRegularSourceSampleDouble(s, n);
    }
  }
  std::vector<unsigned int> transparent;
  for (unsigned int s = 0; s < ns; ++s)
    if (kind[s] == (int32_t)SRC_TRANSPARENT) transparent.push_back(s);
  if (!transparent.empty() && steps) {
    const std::vector<float> &h = m_parameters.getGridIr();
    const size_t m = transparent.size();
    std::vector<float> inf(m * steps), outf(m * steps);
    std::vector<double> ind(single_precision_only ? 0 : m * steps), outd(ind.size());
    for (size_t j = 0; j < m; ++j) {
      std::copy_n(&sf[(size_t)transparent[j] * steps], steps, &inf[j * steps]);
      if (!single_precision_only)
        std::copy_n(&sd[(size_t)transparent[j] * steps], steps, &ind[j * steps]);
    }
    check(pfdtd_transparent_signals(ctx_, (int)m, (int)steps, h.empty() ? nullptr : h.data(),
                                    (int)h.size(), inf.data(), outf.data(),
                                    single_precision_only ? nullptr : ind.data(),
                                    single_precision_only ? nullptr : outd.data()),
          "transparentSources");
    for (size_t j = 0; j < m; ++j) {
      std::copy_n(&outf[j * steps], steps, &sf[(size_t)transparent[j] * steps]);
      if (!single_precision_only)
        std::copy_n(&outd[j * steps], steps, &sd[(size_t)transparent[j] * steps]);
    }
  }
  for (unsigned int r = 0; r < nr; ++r) {
    const nv::Vec3i q = m_parameters.getReceiverElementCoordinates(r);
    rxyz[3 * r] = q.x;
    rxyz[3 * r + 1] = q.y;
    rxyz[3 * r + 2] = q.z;
  }
  check(pfdtd_set_sources(ctx_, (int)ns, sxyz.data(), kind.data(), (int)steps, sf.data(),
                          single_precision_only ? nullptr : sd.data()),
        "setSources");
  check(pfdtd_set_receivers(ctx_, (int)nr, rxyz.data()), "setReceivers");
}

// runSimulation, src/App.cpp:308-348 (launchFDTD3d / launchFDTD3dDouble)
void App::runSimulation() {
  initializeMesh();
  uploadSourcesAndReceivers(!m_mesh.isDouble());
  const unsigned int steps = m_parameters.getNumSteps();
  const unsigned int nr = m_parameters.getNumReceivers();
  int done = 0;
  double ms = 0.0;
  const int prec = m_mesh.isDouble() ? PFDTD_F64 : PFDTD_F32;
  check(pfdtd_run(ctx_, prec, (int)steps, interrupt_tramp, progress_tramp, this, &done, &ms),
        "runSimulation");
  if (m_mesh.isDouble()) {
    responses_double_.assign((size_t)nr * steps, 0.0);
    if (!responses_double_.empty())
      check(pfdtd_get_responses_f64(ctx_, responses_double_.data(), (int64_t)responses_double_.size()),
            "runSimulation");
  } else {
    responses_.assign((size_t)nr * steps, 0.f);
    if (!responses_.empty())
      check(pfdtd_get_responses_f32(ctx_, responses_.data(), (int64_t)responses_.size()),
            "runSimulation");
  }
  time_per_step_ = (float)(ms * 1e-3 / (done > 0 ? done : 1));
}

void App::runVisualization() {
  log_line("runVisualization", "interactive OpenGL visualisation is not part of this library");
  throw(-1);
}

// runCapture, src/App.cpp:350-379: single precision, step by step, captures as scheduled
void App::runCapture() {
  m_mesh.setDouble(false);
  initializeMesh();
  uploadSourcesAndReceivers(true);
  const unsigned int steps = m_parameters.getNumSteps();
  responses_.assign((size_t)steps * m_parameters.getNumReceivers(), 0.f);
  check(pfdtd_begin_steps(ctx_, PFDTD_F32, (int)steps), "runCapture");
  current_step_ = 0;
  if (step_direction_ == 1) {
    // The steps between two scheduled captures advance in one call (the two-level sweep where
    // it applies); results equal the reference's step-by-step loop. Segments also end at every
    // multiple of PROGRESS_MOD (100) steps: the progress callback reports step 100 k there, as
    // launchFDTD3dStep does for those steps (kernels3d.cu:476-479), and the interrupt is
    // polled after every segment -- at most 100 steps apart, where the reference polls after
    // every step (App.cpp:362-366), so an interrupt can let up to 99 more steps complete.
    constexpr unsigned int kProgressMod = 100;
    const auto t0 = std::chrono::steady_clock::now();
    auto seg_t0 = t0;
    unsigned int done = 0;
    while (done < steps) {
      unsigned int stop = std::min(steps, (done / kProgressMod + 1) * kProgressMod);
      for (unsigned int s : step_to_capture_)
        if (s > done && s < stop) stop = s;
      for (unsigned int s : mesh_to_capture_)
        if (s > done && s < stop) stop = s;
      check(pfdtd_steps(ctx_, (int)done, (int)(stop - done)), "runCapture");
      const auto now = std::chrono::steady_clock::now();
      const unsigned int first = done;
      done = stop;
      current_step_ = done;
      captureScheduled();
      if (m_progress && first % kProgressMod == 0)
        m_progress((int)first, (int)steps,
                   (float)(std::chrono::duration<double>(now - seg_t0).count() / (stop - first)));
      seg_t0 = now;
      if (m_interrupt && m_interrupt()) break;
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    time_per_step_ = done ? (float)(sec / done) : 0.f;
  } else {
    float total = 0.f;
    for (unsigned int i = 0; i < steps; ++i) {
      executeStep();
      total += time_per_step_;
      if (m_interrupt && m_interrupt()) break;
    }
    time_per_step_ = steps ? total / (float)steps : 0.f;
  }
  if (!responses_.empty())
    check(pfdtd_get_responses_f32(ctx_, responses_.data(), (int64_t)responses_.size()), "runCapture");
}

// executeStep, src/App.cpp:407-436 (launchFDTD3dStep, captureSliceFast, captureMesh)
void App::executeStep() {
  ensureContext();
  const auto t0 = std::chrono::steady_clock::now();
  const unsigned int step = current_step_;
  check(pfdtd_step(ctx_, (int)step, step_direction_), "executeStep");
  // launchFDTD3dStep reports progress on steps that are multiples of PROGRESS_MOD
  // (kernels3d.cu:476-479)
  if (m_progress && step % 100 == 0)
    m_progress((int)step, (int)m_parameters.getNumSteps(),
               (float)std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  current_step_ += step_direction_;
  captureScheduled();
  time_per_step_ = (float)std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void App::captureScheduled() {
  const unsigned int X = (unsigned int)m_mesh.getDimX(), Y = (unsigned int)m_mesh.getDimY(),
                     Z = (unsigned int)m_mesh.getDimZ();
  for (size_t i = 0; i < step_to_capture_.size(); ++i) {
    if (current_step_ != step_to_capture_[i]) continue;
    const unsigned int orient = slice_orientation_[i];
    const unsigned int slice = slice_to_capture_[i];
    const unsigned int dx = orient == 2 ? Y : X;
    const unsigned int dy = orient == 0 ? Y : Z;
    std::vector<float> p((size_t)dx * dy);
    std::vector<uint8_t> q((size_t)dx * dy);
    if (pfdtd_capture_slice(ctx_, slice, (int)orient, p.data(), q.data()) != PFDTD_OK) {
      log_line("captureSlice", pfdtd_last_error());  // out of range: logged, skipped
      continue;
    }
    saveBitmap(p.data(), q.data(), dx, dy, slice, orient, current_step_);
  }
  for (size_t i = 0; i < mesh_to_capture_.size(); ++i) {
    if (current_step_ != mesh_to_capture_[i]) continue;
    mesh_captures_.emplace_back((size_t)m_mesh.getNumberOfElements());
    check(pfdtd_capture_field(ctx_, mesh_captures_.back().data(), m_mesh.getNumberOfElements()),
          "captureMesh");
  }
}

// resetPressureMesh, src/App.cpp:397-401
void App::resetPressureMesh() {
  ensureContext();
  check(pfdtd_reset_fields(ctx_), "resetPressureMesh");
  current_step_ = 0;
}

// close, src/App.cpp:438-443: release the device resources of this App
void App::close() {
  if (ctx_) pfdtd_ctx_destroy(ctx_);
  ctx_ = nullptr;
  has_domain_ = false;
}

std::vector<float> App::getResponse(unsigned int rec) const {
  std::vector<float> r(m_parameters.getNumSteps(), 0.f);
  for (unsigned int i = 0; i < r.size(); ++i) r[i] = getResponseSampleAt(i, rec);
  return r;
}

std::vector<double> App::getResponseDouble(unsigned int rec) const {
  std::vector<double> r(m_parameters.getNumSteps(), 0.0);
  for (unsigned int i = 0; i < r.size(); ++i) r[i] = getResponseDoubleSampleAt(i, rec);
  return r;
}

// getVolume, src/App.cpp:449-459
float App::getVolume() const {
  const float n = (float)m_mesh.getNumberOfAirElements() + (float)m_mesh.getNumberOfBoundaryElements();
  const float dx = m_parameters.getDx();
  return dx * dx * dx * n;
}

// saveBitmap, src/App.cpp:492-567 and TGAImage::WriteImage (src/io/Image.cpp): 32-bit
// uncompressed TGA, boundary voxels white, pressure as green (positive) / blue (negative)
// on a capture_db dynamic range.
void App::saveBitmap(const float *data, const unsigned char *position_data, unsigned int dim_x,
                     unsigned int dim_y, unsigned int slice, unsigned int orientation,
                     unsigned int step) {
  std::vector<unsigned char> px((size_t)dim_x * dim_y * 4);
  const float dB = capture_db_ / 10.f;
  for (unsigned int i = 0; i < dim_y; ++i) {
    for (unsigned int j = 0; j < dim_x; ++j) {
      const size_t k = (size_t)i * dim_x + j;
      unsigned char r = 0, g = 0, b = 0;
      const unsigned char cp = position_data[k];
      if ((cp >> 7) == 1 && (cp & 0x7F) != 6) {
        r = g = b = 255;
      } else {
        float c = data[k];
        const float sign = c > 0.f ? 1.f : -1.f;
        c = (std::log10(c * c) + dB) / dB;
        c = c > 0 ? c : 0.f;
        c = c < 1.f ? c : 1.f;
        g = (unsigned char)((sign >= 0 ? c : 0.f) * 255);
        b = (unsigned char)((sign < 0 ? c : 0.f) * 255);
      }
      px[4 * k] = b;
      px[4 * k + 1] = g;
      px[4 * k + 2] = r;
      px[4 * k + 3] = 255;
    }
  }
  std::ostringstream name;
  name << capture_dir_ << "/capture_" << orientation << "_" << step << "_" << slice << ".tga";
  std::ofstream o(name.str(), std::ios::out | std::ios::binary);
  const unsigned char hdr[18] = {0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                 (unsigned char)(dim_x & 0xFF), (unsigned char)((dim_x >> 8) & 0xFF),
                                 (unsigned char)(dim_y & 0xFF), (unsigned char)((dim_y >> 8) & 0xFF),
                                 32, 0};
  o.write((const char *)hdr, 18);
  o.write((const char *)px.data(), (std::streamsize)px.size());
}

}  // namespace FDTD
