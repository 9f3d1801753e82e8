// pyfdtd.cpp -- pybind11 module `libPyFDTD`: the reference's Python surface
// (src/AppPy.cpp:96-134) over FDTD::App. Lists or numpy arrays in, numpy arrays out.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>

#include "fdtd_app.h"

namespace py = pybind11;

namespace {

template <typename T>
py::array_t<T> to_numpy(const std::vector<T> &v) {
  py::array_t<T> a(v.size());
  if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(T));
  return a;
}

// The reference throws int -1; surface it as a Python RuntimeError with the same meaning.
template <typename F>
auto guarded(F f) {
  try {
    return f();
  } catch (int code) {
    throw std::runtime_error("FDTD::App failed (" + std::to_string(code) + "), see stderr");
  }
}

}  // namespace

PYBIND11_MODULE(libPyFDTD, m) {
  m.doc() = "ParallelFDTD App on MI355X (SRL_FORWARD core in libpfdtd.so)";
  py::class_<FDTD::App>(m, "App")
      .def(py::init<>())
      .def("initializeDevices", [](FDTD::App &a) { guarded([&] { a.initializeDevices(); }); })
      .def("initializeGeometryFromFile",
           [](FDTD::App &a, const std::string &p) { guarded([&] { a.initializeGeometryFromFile(p); }); })
      // initializeGeometryPy / setLayerIndicesPy, AppPy.cpp:24-50
      .def("initializeGeometryPy",
           [](FDTD::App &a, const std::vector<unsigned int> &idx, const std::vector<float> &vert) {
             a.initializeGeometry(idx.data(), vert.data(), (unsigned int)idx.size(),
                                  (unsigned int)vert.size());
           })
      .def("setLayerIndices",
           [](FDTD::App &a, const std::vector<int> &idx, const std::string &name) {
             a.m_geometry.setLayerIndices(idx, name);
           })
      .def("initializeVoxels",
           [](FDTD::App &a, py::array_t<uint8_t, py::array::c_style | py::array::forcecast> pos,
              py::array_t<uint8_t, py::array::c_style | py::array::forcecast> mat) {
             if (pos.ndim() != 3 || mat.ndim() != 3 || pos.size() != mat.size())
               throw std::invalid_argument("pos and mat must be equal (Z, Y, X) arrays");
             guarded([&] {
               a.initializeVoxels(pos.data(), mat.data(), pos.shape(2), pos.shape(1), pos.shape(0));
             });
           })
      .def("initializeVoxelLabels",
           [](FDTD::App &a, py::array_t<uint8_t, py::array::c_style | py::array::forcecast> lab) {
             if (lab.ndim() != 3) throw std::invalid_argument("labels must be a (Z, Y, X) array");
             guarded([&] { a.initializeVoxelLabels(lab.data(), lab.shape(2), lab.shape(1), lab.shape(0)); });
           })
      .def("addSource", &FDTD::App::addSource)
      // addSourceDataFloat / addSourceDataDouble, AppPy.cpp:66-94: one vector of
      // num_steps*num_sources samples added as a single input-data entry
      .def("addSourceDataFloat",
           [](FDTD::App &a, const std::vector<float> &d, int num_steps, int num_sources) {
             std::vector<float> v((size_t)num_steps * num_sources);
             for (size_t i = 0; i < v.size(); ++i) v[i] = d.at(i);
             a.addSourceData(v);
           })
      .def("addSourceDataDouble",
           [](FDTD::App &a, const std::vector<double> &d, int num_steps, int num_sources) {
             std::vector<double> v((size_t)num_steps * num_sources);
             for (size_t i = 0; i < v.size(); ++i) v[i] = d.at(i);
             a.addSourceDataDouble(v);
           })
      .def("addReceiver", &FDTD::App::addReceiver)
      // addSurfaceMaterials, AppPy.cpp:52-64
      .def("addSurfaceMaterials",
           [](FDTD::App &a, const std::vector<float> &coefs, unsigned int n_surf, unsigned int n_coef) {
             if (coefs.size() < (size_t)n_surf * n_coef)
               throw std::invalid_argument("material list shorter than surfaces x coefficients");
             a.m_materials.addMaterials(coefs.data(), n_surf, n_coef);
           })
      .def("setSpatialFs", &FDTD::App::setSpatialFs)
      .def("setNumSteps", &FDTD::App::setNumSteps)
      .def("setUpdateType", &FDTD::App::setUpdateType)
      .def("setOctave", [](FDTD::App &a, unsigned int o) { a.m_parameters.setOctave(o); })
      .def("setUniform", &FDTD::App::setUniformMaterial)
      .def("setUniformMaterial", &FDTD::App::setUniformMaterial)
      .def("runVisualization", [](FDTD::App &a) { guarded([&] { a.runVisualization(); }); })
      .def("runSimulation", [](FDTD::App &a) { guarded([&] { a.runSimulation(); }); })
      .def("runCapture", [](FDTD::App &a) { guarded([&] { a.runCapture(); }); })
      .def("executeStep", [](FDTD::App &a) { guarded([&] { a.executeStep(); }); })
      .def("resetPressureMesh", [](FDTD::App &a) { guarded([&] { a.resetPressureMesh(); }); })
      .def("invertTime", &FDTD::App::invertTime)
      .def("getResponse", [](const FDTD::App &a, unsigned int r) { return to_numpy(a.getResponse(r)); })
      .def("getResponseDouble",
           [](const FDTD::App &a, unsigned int r) { return to_numpy(a.getResponseDouble(r)); })
      .def("getResponseSampleAt", &FDTD::App::getResponseSampleAt)
      .def("getResponseDoubleSampleAt", &FDTD::App::getResponseDoubleSampleAt)
      .def("getResponseSize", &FDTD::App::getResponseSize)
      .def("forcePartitionTo", &FDTD::App::setForcePartitionTo)
      .def("addSliceToCapture", &FDTD::App::addSliceToCapture)
      .def("addMeshToCapture", &FDTD::App::addMeshToCapture)
      .def("getNumberOfMeshCaptures", &FDTD::App::getNumberOfMeshCaptures)
      .def("getMeshCaptureAt",
           [](FDTD::App &a, unsigned int i) {
             const float *p = a.getMeshCaptureAt(i);
             py::array_t<float> out({(py::ssize_t)a.m_mesh.getDimZ(), (py::ssize_t)a.m_mesh.getDimY(),
                                     (py::ssize_t)a.m_mesh.getDimX()});
             std::memcpy(out.mutable_data(), p, sizeof(float) * (size_t)a.m_mesh.getNumberOfElements());
             return out;
           })
      .def("setDouble", &FDTD::App::setDouble)
      .def("setCapturedB", &FDTD::App::setCapturedB)
      .def("setCaptureDirectory", &FDTD::App::setCaptureDirectory)
      .def("setReferenceQuirks", &FDTD::App::setReferenceQuirks)
      .def("readGridIr", [](FDTD::App &a, const std::string &p) { a.m_parameters.readGridIr(p); })
      .def("close", &FDTD::App::close)
      .def("getMvox", &FDTD::App::getMvoxPerSec)
      .def("getNumElems", &FDTD::App::getNumElements)
      .def("getTimePerStep", &FDTD::App::getTimePerStep)
      .def("getVolume", &FDTD::App::getVolume)
      .def("getDx", [](const FDTD::App &a) { return a.m_parameters.getDx(); })
      // SimulationParameters accessors (src/base/SimulationParameters.h:128-176)
      .def("getSourceSample",
           [](const FDTD::App &a, unsigned int i, unsigned int n) { return a.m_parameters.getSourceSample(i, n); })
      .def("getSourceSampleDouble",
           [](const FDTD::App &a, unsigned int i, unsigned int n) {
             return a.m_parameters.getSourceSampleDouble(i, n);
           })
      .def("getSourceElementCoordinates",
           [](const FDTD::App &a, unsigned int i) {
             const nv::Vec3i q = a.m_parameters.getSourceElementCoordinates(i);
             return py::make_tuple(q.x, q.y, q.z);
           })
      .def("getReceiverElementCoordinates",
           [](const FDTD::App &a, unsigned int i) {
             const nv::Vec3i q = a.m_parameters.getReceiverElementCoordinates(i);
             return py::make_tuple(q.x, q.y, q.z);
           })
      .def("getNumSources", [](const FDTD::App &a) { return a.m_parameters.getNumSources(); })
      .def("getNumReceivers", [](const FDTD::App &a) { return a.m_parameters.getNumReceivers(); })
      .def("setGridIr", [](FDTD::App &a, const std::vector<float> &h) { a.m_parameters.setGridIr(h); })
      .def("getParameterVector", [](const FDTD::App &a) { return a.m_parameters.getParameterVector(); })
      .def("getParameterVectorDouble",
           [](const FDTD::App &a) { return a.m_parameters.getParameterVectorDouble(); })
      .def("getDims", [](const FDTD::App &a) {
        return py::make_tuple(a.m_mesh.getDimX(), a.m_mesh.getDimY(), a.m_mesh.getDimZ());
      });
}
