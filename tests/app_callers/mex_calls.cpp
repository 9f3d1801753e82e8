// Every FDTD::App call the reference's MEX gateway makes (matlab/mex_FDTD.cpp:183-365), with
// the argument types it passes, through the header path it includes (matlab/mex_includes.h:5).
// MATLAB's mex.h is not needed: the mx* inputs are replaced by plain arrays of the same types.
// Compiled (not run) by tests/test_app_callers.py against the facade headers.
#include <signal.h>

#include <cstdio>
#include <vector>

#include "./includes/App.h"

FDTD::App app;

extern "C" {
void progressCallback(int step, int max_step, float t_per_step) {
  std::printf("Step %d/%d, %f s per step, %f s left\n", step, max_step, t_per_step,
              t_per_step * (float)(max_step - step));
}
bool interruptPending() { return false; }
}

void sigAbortHandler(int signum) {
  if (signum == SIGINT) app.close();
}

int gateway(float *vertices, unsigned int size_of_vertices, unsigned int *indices,
            unsigned int size_of_indices, float *materials, unsigned int number_of_surfaces,
            unsigned int number_of_coefficients, float *src_list, unsigned int number_of_sources,
            float *rec_list, unsigned int number_of_receivers, double *input_data,
            unsigned int number_of_input_data_vectors, unsigned int number_of_input_samples,
            unsigned int spatial_fs, unsigned int number_of_steps, unsigned int update_type,
            unsigned int visualization, unsigned int *capture_instructions,
            unsigned int number_of_captures, unsigned int *mesh_capture_instructions,
            unsigned int number_of_mesh_captures, unsigned int *double_precision,
            unsigned int *force_partition_to, unsigned int *octave, float *out,
            double *out_double, float *out_mesh) {
  loggerInit();
  log_msg<LOG_INFO>(L"Mex: begin");
  signal(SIGINT, sigAbortHandler);
  app.initializeDevices();
  app.m_interrupt = (InterruptCallback)interruptPending;
  app.m_progress = (ProgressCallback)progressCallback;
  app.initializeGeometry(indices, vertices, size_of_indices, size_of_vertices);
  app.m_materials.addMaterials(materials, number_of_surfaces, number_of_coefficients);
  app.m_parameters.setSpatialFs(spatial_fs);
  app.m_parameters.setNumSteps(number_of_steps);
  app.m_parameters.readGridIr("./Data/grid_ir.txt");
  app.m_parameters.setUpdateType((enum UpdateType)update_type);
  app.m_parameters.setOctave(*octave);
  app.setForcePartitionTo((int)(*force_partition_to));
  for (unsigned int i = 0; i < number_of_input_data_vectors; i++) {
    if (*double_precision == 1) {
      std::vector<double> v(number_of_input_samples, 0.0);
      for (unsigned int j = 0; j < number_of_input_samples; j++)
        v.at(j) = input_data[j + i * number_of_input_samples];
      app.m_parameters.addInputDataDouble(v);
    } else {
      std::vector<float> v(number_of_input_samples, 0.f);
      for (unsigned int j = 0; j < number_of_input_samples; j++)
        v.at(j) = (float)input_data[j + i * number_of_input_samples];
      app.m_parameters.addInputData(v);
    }
  }
  for (unsigned int i = 0; i < number_of_sources; i++) {
    const unsigned int k = i * 6;
    app.m_parameters.addSource(Source(src_list[k], src_list[k + 1], src_list[k + 2],
                                      (enum SrcType)((unsigned int)src_list[k + 3]),
                                      (enum InputType)((unsigned int)src_list[k + 4]),
                                      src_list[k + 5]));
  }
  for (unsigned int i = 0; i < number_of_receivers; i++) {
    const unsigned int k = i * 3;
    app.m_parameters.addReceiver(Receiver(rec_list[k], rec_list[k + 1], rec_list[k + 2]));
  }
  for (unsigned int i = 0; i < number_of_captures; i++)
    app.addSliceToCapture(capture_instructions[i * 3], capture_instructions[i * 3 + 1],
                          capture_instructions[i * 3 + 2]);
  for (unsigned int i = 0; i < number_of_mesh_captures; i++)
    app.addMeshToCapture(mesh_capture_instructions[i]);
  try {
    if (visualization) {
      app.runVisualization();
    } else if (number_of_captures == 0 && number_of_mesh_captures == 0) {
      if (*double_precision == 1) app.m_mesh.setDouble(true);
      app.runSimulation();
    } else {
      app.runCapture();
    }
  } catch (...) {
    app.close();
    return -1;
  }
  const unsigned int n_steps = app.m_parameters.getNumSteps();
  const unsigned int n_mesh_captures = app.getNumberOfMeshCaptures();
  const unsigned int n_elements = app.m_mesh.getNumberOfElements();
  float stats[6] = {(float)app.getNumElements(), app.getTimePerStep(),
                    (float)app.m_mesh.getDimX(), (float)app.m_mesh.getDimY(),
                    (float)app.m_mesh.getDimZ(), app.m_parameters.getDx()};
  (void)stats;
  if (app.getResponseSize() != 0) {
    for (unsigned int i = 0; i < n_steps; i++)
      for (unsigned int j = 0; j < number_of_receivers; j++) {
        if (app.m_mesh.isDouble())
          out_double[i * number_of_receivers + j] = app.getResponseDoubleSampleAt(i, j);
        else
          out[i * number_of_receivers + j] = app.getResponseSampleAt(i, j);
      }
  }
  for (unsigned int i = 0; i < n_mesh_captures; i++) {
    float *mesh_capture = app.getMeshCaptureAt(i);
    for (unsigned int j = 0; j < n_elements; j++) out_mesh[j * n_mesh_captures + i] = *(mesh_capture + j);
  }
  app.close();
  return 0;
}
