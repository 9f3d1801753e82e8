// fdtd_app.h -- the FDTD::App host facade over the C ABI (include/pfdtd.h).
//
// Same class and method names, argument meaning and error behaviour as the reference's
// application layer (src/App.h:50-396, src/base/SimulationParameters.h, src/base/SrcRec.h,
// src/base/MaterialHandler.h), so the Python binding, a MEX gateway or a C++ main call it
// unchanged. The header pulls in no HIP header: callers compile with a plain C++ compiler
// and link libpfdtd_app.so (which links libpfdtd.so).
//
// Errors follow the reference: device/runtime failures and unusable inputs throw int -1
// (cudasafe, src/kernels/cudaUtils.h:47-51); out-of-range indices throw std::out_of_range.
//
// The domain comes in as a closed triangle mesh (initializeGeometry / initializeGeometryPy /
// initializeGeometryFromFile, voxelized on the device by pfdtd_voxelize at every
// initializeMesh, as the reference voxelizes in App::initializeMesh) or directly as voxels
// (initializeVoxels / initializeVoxelLabels); runSimulation without either throws -1.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "global_includes.h"  // logger, reflection2Admitance / admitance2Reflection

struct pfdtd_ctx;

// enum SrcType / InputType, src/base/SrcRec.h:33-34
enum SrcType { SRC_HARD, SRC_SOFT, SRC_TRANSPARENT };
enum InputType { IMPULSE, GAUSSIAN, SINE, DATA };
// enum UpdateType, src/base/SimulationParameters.h:39
enum UpdateType { SRL_FORWARD, SHARED, SRL };

typedef bool (*InterruptCallback)(void);
typedef void (*ProgressCallback)(int, int, float);

namespace nv {
struct Vec3f {
  float x = 0.f, y = 0.f, z = 0.f;
  Vec3f() = default;
  Vec3f(float a, float b, float c) : x(a), y(b), z(c) {}
};
struct Vec3i {
  int x = 0, y = 0, z = 0;
  Vec3i() = default;
  Vec3i(int a, int b, int c) : x(a), y(b), z(c) {}
};
}  // namespace nv


// Position / Source / Receiver, src/base/SrcRec.h
class Position {
 public:
  Position() = default;
  Position(float x, float y, float z) : p_(x, y, z) {}
  virtual ~Position() = default;
  nv::Vec3f getP() const { return p_; }
  // Position::getElementIdx, SrcRec.cpp:26-34 (nv::ROUND = floor(x + 0.5f))
  nv::Vec3i getElementIdx(unsigned fs, float c, float lambda) const;

 protected:
  nv::Vec3f p_;
};

class Source : public Position {
 public:
  Source() = default;
  Source(float x, float y, float z) : Position(x, y, z) {}
  Source(float x, float y, float z, SrcType t) : Position(x, y, z), source_type_(t) {}
  Source(float x, float y, float z, SrcType t, InputType in, unsigned int data_idx)
      : Position(x, y, z), source_type_(t), input_type_(in), input_data_idx_(data_idx) {}
  SrcType getSourceType() const { return source_type_; }
  InputType getInputType() const { return input_type_; }
  unsigned int getInputDataIdx() const { return input_data_idx_; }

 private:
  SrcType source_type_ = SRC_HARD;
  InputType input_type_ = IMPULSE;
  unsigned int input_data_idx_ = 0;
};

class Receiver : public Position {
 public:
  Receiver() = default;
  Receiver(float x, float y, float z) : Position(x, y, z) {}
};

// SimulationParameters, src/base/SimulationParameters.{h,cpp}
class SimulationParameters {
 public:
  void readGridIr(const std::string &path);
  void setGridIr(const std::vector<float> &h) { grid_ir_ = h; }
  void setUpdateType(UpdateType t);
  void setC(float c) { c_ = c; }
  void setLambda(double l) { lambda_ = l; }
  void setOctave(unsigned int o) { octave_ = o; }
  void setNumSteps(unsigned int n) { num_steps_ = n; }
  void setSpatialFs(unsigned int fs) { spatial_fs_ = fs; }
  void setAddPaddingToElementIdx(bool b) { add_padding_to_element_idx_ = b; }

  UpdateType getUpdateType() const { return update_type_; }
  float getC() const { return c_; }
  double getLambda() const { return lambda_; }
  float getDx() const;
  unsigned int getOctave() const { return octave_; }
  unsigned int getNumSteps() const { return num_steps_; }
  unsigned int getSpatialFs() const { return spatial_fs_; }

  void addSource(float x, float y, float z) { sources_.push_back(Source(x, y, z)); }
  void addSource(const Source &s) { sources_.push_back(s); }
  void addReceiver(float x, float y, float z) { receivers_.push_back(Receiver(x, y, z)); }
  void addReceiver(const Receiver &r) { receivers_.push_back(r); }
  void removeSource(unsigned int i);
  void removeReceiver(unsigned int i);
  void updateSourceAt(unsigned int i, const Source &s) { sources_.at(i) = s; }
  void updateReceiverAt(unsigned int i, const Receiver &r) { receivers_.at(i) = r; }
  void resetSourcesAndReceivers() {
    sources_.clear();
    receivers_.clear();
  }
  void addInputData(const float *data, unsigned int n);
  void addInputData(const std::vector<float> &d) { input_f_.push_back(d); }
  void addInputDataDouble(const std::vector<double> &d) { input_d_.push_back(d); }

  float getSourceSample(unsigned int src, unsigned int step) const;
  double getSourceSampleDouble(unsigned int src, unsigned int step) const;
  float getInputDataSample(unsigned int idx, unsigned int sample) const;
  double getInputDataSampleDouble(unsigned int idx, unsigned int sample) const;
  float getGridIrDataSample(unsigned int sample) const;
  nv::Vec3i getSourceElementCoordinates(unsigned int i) const;
  nv::Vec3i getReceiverElementCoordinates(unsigned int i) const;
  unsigned int getNumSources() const { return (unsigned int)sources_.size(); }
  unsigned int getNumReceivers() const { return (unsigned int)receivers_.size(); }
  Source getSource(unsigned int i) const { return sources_.at(i); }
  Receiver getReceiver(unsigned int i) const { return receivers_.at(i); }
  // getRegularSourceSample(+Double), SimulationParameters.cpp:262-298, :312-348
  float getRegularSourceSample(unsigned int src, unsigned int step) const {
    return regularSample(src, step);
  }
  double getRegularSourceSampleDouble(unsigned int src, unsigned int step) const {
    return regularSampleDouble(src, step);
  }
  const std::vector<float> &getGridIr() const { return grid_ir_; }
  std::vector<float> getParameterVector() const;        // getParameterPtr, :379-387
  std::vector<double> getParameterVectorDouble() const;  // getParameterPtrDouble, :389-396

 private:
  float regularSample(unsigned int src, unsigned int step) const;
  double regularSampleDouble(unsigned int src, unsigned int step) const;

  UpdateType update_type_ = SRL_FORWARD;
  float c_ = 344.f;
  double lambda_ = 1.0 / std::sqrt(3.0);
  unsigned int octave_ = 0;
  unsigned int num_steps_ = 1;
  unsigned int spatial_fs_ = 7000;
  bool add_padding_to_element_idx_ = true;
  std::vector<Source> sources_;
  std::vector<Receiver> receivers_;
  std::vector<std::vector<float>> input_f_;
  std::vector<std::vector<double>> input_d_;
  std::vector<float> grid_ir_;
};

// MaterialHandler, src/base/MaterialHandler.{h,cpp}: per-surface material indices and a
// deduplicated table of 20 coefficients per unique material.
class MaterialHandler {
 public:
  static constexpr unsigned int kCoefNum = 20;  // MATERIAL_COEF_NUM
  void addMaterials(const float *coefs, unsigned int n_surfaces, unsigned int n_coefs);
  unsigned int addSurfaceMaterial(std::vector<float> coefs);
  void setGlobalMaterial(unsigned int n_surfaces, float coef);
  void coefsAreAdmittances() { admittance_ = true; }
  void coefsAreReflectances() { admittance_ = false; }
  unsigned int getNumberOfCoefficients() const { return kCoefNum; }
  unsigned int getNumberOfSurfaces() const { return (unsigned int)indices_.size(); }
  unsigned int getNumberOfUniqueMaterials() const { return (unsigned int)unique_.size(); }
  unsigned int getMaterialIdxAt(unsigned int s) const { return indices_.at(s); }
  float getUniqueCoefAt(unsigned int m, unsigned int c) const;
  float getSurfaceCoefAt(unsigned int s, unsigned int c) const;
  std::vector<float> getMaterialCoefficients() const;  // getMaterialCoefficientPtr layout
  float getMeanAbsorption(unsigned int octave) const;
  void setMaterialIndexAt(unsigned int s, unsigned char m) {
    if (s < indices_.size()) indices_[s] = m;
  }

 private:
  bool admittance_ = true;
  std::vector<unsigned char> indices_;
  std::vector<std::vector<float>> unique_;
};

// GeometryHandler (triangle soup kept for API compatibility and statistics).
class GeometryHandler {
 public:
  void initialize(const unsigned int *indices, const float *vertices, unsigned int n_idx,
                  unsigned int n_vert);
  void initialize(const std::vector<unsigned int> &idx, const std::vector<float> &vert) {
    initialize(idx.data(), vert.data(), (unsigned int)idx.size(), (unsigned int)vert.size());
  }
  void setLayerIndices(const std::vector<int> &idx, const std::string &name);
  unsigned int getNumberOfTriangles() const { return (unsigned int)(indices_.size() / 3); }
  unsigned int getNumberOfVertices() const { return (unsigned int)(vertices_.size() / 3); }
  const std::vector<float> &vertices() const { return vertices_; }
  const std::vector<unsigned int> &indices() const { return indices_; }
  const std::vector<std::pair<std::string, std::vector<int>>> &layers() const { return layers_; }

 private:
  std::vector<float> vertices_;
  std::vector<unsigned int> indices_;
  std::vector<std::pair<std::string, std::vector<int>>> layers_;
};

// The voxel domain on the device (the CudaMesh accessors callers use).
class VoxelMesh {
 public:
  int getDimX() const { return (int)X_; }
  int getDimY() const { return (int)Y_; }
  int getDimZ() const { return (int)Z_; }
  int64_t getNumberOfElements() const { return X_ * Y_ * Z_; }
  uint64_t getNumberOfAirElements() const { return air_; }
  uint64_t getNumberOfBoundaryElements() const { return boundary_; }
  bool isDouble() const { return double_; }
  void setDouble(bool d) { double_ = d; }

 private:
  int64_t X_ = 0, Y_ = 0, Z_ = 0;
  uint64_t air_ = 0, boundary_ = 0;
  bool double_ = false;

 public:
  void set(int64_t X, int64_t Y, int64_t Z, uint64_t air, uint64_t bnd) {
    X_ = X;
    Y_ = Y;
    Z_ = Z;
    air_ = air;
    boundary_ = bnd;
  }
};

namespace FDTD {

class App {
 public:
  App();
  ~App();
  App(const App &) = delete;
  App &operator=(const App &) = delete;

  SimulationParameters m_parameters;
  GeometryHandler m_geometry;
  MaterialHandler m_materials;
  VoxelMesh m_mesh;
  InterruptCallback m_interrupt;
  ProgressCallback m_progress;

  // ---- setup (App.h:80-116)
  void initializeDevices();
  void initializeGeometry(const unsigned int *indices, const float *vertices,
                          unsigned int number_of_indices, unsigned int number_of_vertices);
  // Legacy-VTK polydata (FileReader::readVTK, src/io/FileReader.cpp:41-89); throws -1 on a
  // file that cannot be read (App.cpp:133-139).
  void initializeGeometryFromFile(const std::string &path);
  void setupDefaultCallbacks();
  // Voxel domain: pre-classified position/material bytes (layout (z*Y+y)*X+x) ...
  void initializeVoxels(const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y, int64_t Z);
  // ... or a label grid (0 inside, m+1 solid of material m) classified on the device.
  void initializeVoxelLabels(const uint8_t *labels, int64_t X, int64_t Y, int64_t Z);

  // ---- run (App.h:118-147)
  void runSimulation();
  void runVisualization();  // interactive GL: not part of this library, throws -1
  void runCapture();
  void executeStep();
  void resetPressureMesh();
  void invertTime() { step_direction_ *= -1; }
  void close();

  // ---- captures and getters (App.h:148-247)
  void setCapturedB(float db) { capture_db_ = db; }
  void setCaptureDirectory(const std::string &dir) { capture_dir_ = dir; }
  void addSliceToCapture(unsigned int slice, unsigned int step, unsigned int orientation) {
    step_to_capture_.push_back(step);
    slice_to_capture_.push_back(slice);
    slice_orientation_.push_back(orientation);
  }
  void addMeshToCapture(unsigned int step) { mesh_to_capture_.push_back(step); }
  void saveBitmap(const float *data, const unsigned char *position_data, unsigned int dim_x,
                  unsigned int dim_y, unsigned int slice, unsigned int orientation,
                  unsigned int step);
  float getTimePerStep() const { return time_per_step_; }
  unsigned int getNumElements() const { return (unsigned int)m_mesh.getNumberOfElements(); }
  unsigned int getResponseSize() const {
    return (unsigned int)(m_mesh.isDouble() ? responses_double_.size() : responses_.size());
  }
  float getMvoxPerSec() const {
    return (float)((1.f / time_per_step_ * (double)m_mesh.getNumberOfElements()) / 1e6);
  }
  float *getResponsePointer() { return responses_.data(); }
  float getResponseSampleAt(unsigned int step, unsigned int rec) const {
    return responses_.at((size_t)m_parameters.getNumSteps() * rec + step);
  }
  double getResponseDoubleSampleAt(unsigned int step, unsigned int rec) const {
    return responses_double_.at((size_t)m_parameters.getNumSteps() * rec + step);
  }
  std::vector<float> getResponse(unsigned int rec) const;
  std::vector<double> getResponseDouble(unsigned int rec) const;
  float *getMeshCaptureAt(unsigned int i) { return mesh_captures_.at(i).data(); }
  unsigned int getNumberOfMeshCaptures() const { return (unsigned int)mesh_captures_.size(); }
  float getVolume() const;

  // ---- reference Python/MEX convenience setters (App.h:345-394)
  void addSource(float x, float y, float z, int type, int signal, int input_signal_idx) {
    m_parameters.addSource(Source(x, y, z, (SrcType)type, (InputType)signal,
                                  (unsigned int)input_signal_idx));
  }
  void addReceiver(float x, float y, float z) { m_parameters.addReceiver(x, y, z); }
  void setSpatialFs(unsigned int fs) { m_parameters.setSpatialFs(fs); }
  void setNumSteps(unsigned int n) { m_parameters.setNumSteps(n); }
  void setUpdateType(int i) { m_parameters.setUpdateType((UpdateType)i); }
  void setUniformMaterial(float R) {
    m_materials.setGlobalMaterial(m_geometry.getNumberOfTriangles() ? m_geometry.getNumberOfTriangles() : 1,
                                  reflection2Admitance(R));
  }
  void setDouble(bool d) { m_mesh.setDouble(d); }
  void setForcePartitionTo(int n) { force_partition_to_ = n; }  // one GPU holds the domain
  void addSourceData(const std::vector<float> &d) { m_parameters.addInputData(d); }
  void addSourceDataDouble(const std::vector<double> &d) { m_parameters.addInputDataDouble(d); }
  // Reproduce the reference's shipped behaviour instead of its intent: material slot
  // mat*20*octave (kernels3d.cu:513) and/or soft sources that overwrite (cudaMesh.h:362-366).
  void setReferenceQuirks(bool literal_material_slot, bool soft_source_overwrites) {
    quirk_flags_ = (literal_material_slot ? 1 : 0) | (soft_source_overwrites ? 2 : 0);
  }

  unsigned int current_step_ = 0;
  int step_direction_ = 1;

 private:
  void ensureContext();
  void initializeMesh();
  void uploadSourcesAndReceivers(bool single_precision_only);
  void captureScheduled();  // the slice and mesh captures scheduled at current_step_
  void check(int rc, const char *what);

  pfdtd_ctx *ctx_ = nullptr;
  int device_ = 0;
  bool has_domain_ = false;
  bool geometry_domain_ = false;  // the domain is voxelized from m_geometry at initializeMesh
  std::vector<uint8_t> labels_;  // label grid of initializeVoxelLabels (reclassification)
  int64_t labels_dims_[3] = {0, 0, 0};
  int domain_type_ = 0;          // update type the resident domain was classified for
  int force_partition_to_ = -1;
  int quirk_flags_ = 0;
  float capture_db_ = 60.f;
  std::string capture_dir_ = ".";
  float time_per_step_ = 0.f;
  std::vector<unsigned int> step_to_capture_, slice_to_capture_, slice_orientation_;
  std::vector<unsigned int> mesh_to_capture_;
  std::vector<float> responses_;
  std::vector<double> responses_double_;
  std::vector<std::vector<float>> mesh_captures_;
};

}  // namespace FDTD
