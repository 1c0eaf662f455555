// lgpu_setup.h -- set-up kernels of the reference's Sedov problem
// (Domain::Domain, lulesh-init.cc:16-194, and the set-up routines it calls),
// run by lgpu_create_sedov.
#pragma once

#include <hip/hip_runtime.h>

#include "lgpu_kernels.h"

namespace lgpu {

struct SedovSetup {
   int32_t size;              // elements per edge
   const int32_t *runStart;   // device: first element of each region run, numRuns + 1 entries
   const int32_t *runDest;    // device: work-list position of each run's first element
   int32_t numRuns;
};

// Fills the mesh, connectivity, boundary, work-list and initial-state arrays
// of s (an arena zeroed beforehand) on stream st.  Enqueues only.
hipError_t launch_sedov_setup(const DevState &s, const SedovSetup &t, hipStream_t st);

} // namespace lgpu
