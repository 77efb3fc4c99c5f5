// Forwarding header: lets the reference's `#include "classical_spin/gpu/mixed_lattice_gpu_api.h"`
// (mixed_lattice.h GPU glue) resolve to the libcsl_hip.so drop-in API.
#pragma once
#include "classical_spin_amd/mixed_lattice_gpu_api.h"
