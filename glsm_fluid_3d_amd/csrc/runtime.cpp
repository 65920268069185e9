// runtime.cpp -- the library stream and HIP error reporting.
#include <mutex>
#include <string>

#include "glsm_internal.hpp"

namespace glsm {
namespace {
std::mutex g_mu;
hipStream_t g_stream = nullptr;
int g_device = -1;
}  // namespace

void set_device(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_stream && device != g_device)
    throw std::runtime_error("glsm: the library stream already lives on device " + std::to_string(g_device));
  GLSM_HIP(hipSetDevice(device));
  g_device = device;
}

// One stream for every Gel/Coupler/Fluid: the reference's per-object streams
// (gel.cu:674, fluid.cu:243, coupling.cu:127) are unordered with respect to
// each other, while its drivers assume program order.  Default priority,
// blocking with respect to the null stream, so a driver's plain hipMemcpy
// sees all queued work.
hipStream_t stream() {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_stream) {
    if (g_device < 0) GLSM_HIP(hipGetDevice(&g_device));
    GLSM_HIP(hipStreamCreate(&g_stream));
  }
  return g_stream;
}

void throw_hip(hipError_t e, const char* what, const char* file, int line) {
  throw std::runtime_error(std::string(file) + ":" + std::to_string(line) + ": " + what + " failed: " +
                           hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
}

}  // namespace glsm
