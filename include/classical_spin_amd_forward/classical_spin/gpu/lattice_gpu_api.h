// Forwarding header: lets the reference's `#include "classical_spin/gpu/lattice_gpu_api.h"`
// (lattice.h GPU glue) resolve to the libcsl_hip.so drop-in API.
#pragma once
#include "classical_spin_amd/lattice_gpu_api.h"
