// device.cpp -- see device.h.
#include "device.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <unordered_map>

#include "planner.h"

namespace hpttx {

bool device_arch(std::string &arch, std::string &err) {
  static std::mutex mu;
  static std::unordered_map<int, std::string> names;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    err = std::string("no HIP device: ") + hipGetErrorString(e);
    return false;
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = names.find(dev);
    if (it != names.end()) {
      arch = it->second;
      return true;
    }
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    err = std::string("device properties: ") + hipGetErrorString(e);
    return false;
  }
  std::string name(prop.gcnArchName);
  const size_t colon = name.find(':');
  if (colon != std::string::npos) name.resize(colon);
  std::lock_guard<std::mutex> lk(mu);
  names[dev] = name;
  arch = name;
  return true;
}

int require_target_device(std::string &err) {
  std::string arch;
  if (!device_arch(arch, err)) return kEHip;
  if (arch != kTargetArch) {
    err = "device is " + arch + "; this library's kernels are built for " + kTargetArch;
    return kEUnsupported;
  }
  return kOk;
}

}  // namespace hpttx
