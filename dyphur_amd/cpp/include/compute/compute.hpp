#pragma once
// Everything of the compute shim (compute/include/compute/compute.hpp).
// kernel.hpp and atomics.hpp need a HIP compiler; host-only code includes
// the other headers directly.
#include "device.hpp"
#include "stream.hpp"
#include "buffer.hpp"
#include "reduction.hpp"
#include "sort.hpp"
#include "kernel.hpp"
#include "atomics.hpp"
