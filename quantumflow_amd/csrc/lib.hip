// Library-level entry points of the qfx C-ABI: errors, version, init.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

#include "qfx_common.h"

namespace qfx {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int check_hip(hipError_t err, const char* what) {
  if (err == hipSuccess) return QFX_OK;
  return set_error(QFX_EHIP, std::string(what) + ": " + hipGetErrorString(err));
}

int check_launch(const char* what) { return check_hip(hipGetLastError(), what); }

int check_targets(int n, int k, const int* targets, int kmax, const char* fn) {
  if (n < 1 || n > QFX_MAX_QUBITS)
    return set_error(QFX_ERANGE, std::string(fn) + ": qubit count out of range");
  if (k < 1 || k > kmax || k > n)
    return set_error(QFX_ERANGE, std::string(fn) + ": gate width out of range");
  if (!targets) return set_error(QFX_EINVAL, std::string(fn) + ": null targets");
  for (int j = 0; j < k; ++j) {
    if (targets[j] < 0 || targets[j] >= n)
      return set_error(QFX_EINVAL, std::string(fn) + ": target bit out of range");
    for (int i = 0; i < j; ++i)
      if (targets[i] == targets[j])
        return set_error(QFX_EINVAL, std::string(fn) + ": repeated target qubit");
  }
  return QFX_OK;
}

void sort_bits(int k, const int* targets, int* sorted) {
  for (int j = 0; j < k; ++j) sorted[j] = targets[j];
  std::sort(sorted, sorted + k);
}

}  // namespace qfx

extern "C" {

const char* qfx_last_error(void) { return qfx::g_last_error.c_str(); }

const char* qfx_version(void) { return "qfx 0.1.0 gfx950"; }

int qfx_init(int64_t* props) {
  int dev = 0;
  int rc = qfx::check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc == QFX_EHIP ? qfx::set_error(QFX_ENODEV, "no HIP device") : rc;
  hipDeviceProp_t p;
  rc = qfx::check_hip(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
  if (rc) return rc;
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
    return qfx::set_error(QFX_ENODEV, std::string("device is ") + p.gcnArchName +
                                          ", this build targets gfx950 only");
  if (props) {
    props[0] = p.multiProcessorCount;
    props[1] = (int64_t)p.sharedMemPerBlock;
    props[2] = p.warpSize;
    props[3] = p.l2CacheSize;
    props[4] = p.clockRate;
    props[5] = (int64_t)(p.totalGlobalMem >> 20);
    props[6] = p.maxThreadsPerBlock;
    int wall_khz = 0;
    rc = qfx::check_hip(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev),
                        "hipDeviceGetAttribute");
    if (rc) return rc;
    props[7] = wall_khz;
  }
  return QFX_OK;
}

}  // extern "C"
