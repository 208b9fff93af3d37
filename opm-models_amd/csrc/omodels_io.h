// omodels_io.h -- restart and VTK files (host code, io.cpp).
#pragma once

#include <stdint.h>

namespace omodels {

// Restart file of a structured nx x ny x nz grid: primaries [N][3], gas
// meaning per cell, the simulator's time and time-step index.
void writeRestart(const char* path, long nx, long ny, long nz, const double* pv, const uint8_t* meaning, double time,
                  int timeStepIdx);
void readRestart(const char* path, long nx, long ny, long nz, double* pv, uint8_t* meaning, double* time,
                 int* timeStepIdx);

// VTK unstructured-grid file of cell fields [nfields][N] on the hexahedra of the grid.
void writeVtu(const char* path, long nx, long ny, long nz, double dx, double dy, double dz, int nfields,
              const char* const* names, const double* fields);

}  // namespace omodels
