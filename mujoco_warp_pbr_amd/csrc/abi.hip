// abi.hip — library-level entry points of the C ABI (version, errors, struct sizes).
#include "abi_util.hpp"

namespace mjpbr {

static thread_local char g_error[512] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace mjpbr

#ifndef MJPBR_HEADER_HASH
#define MJPBR_HEADER_HASH "unknown"
#endif

extern "C" int mjpbr_abi_version(void) { return MJPBR_ABI_VERSION; }

extern "C" const char* mjpbr_header_hash(void) { return MJPBR_HEADER_HASH; }

extern "C" const char* mjpbr_last_error(void) { return mjpbr::g_error; }

extern "C" int mjpbr_struct_sizes(int64_t* sizes, int n) {
  const int64_t all[] = {
      (int64_t)sizeof(MjpbrGeometry), (int64_t)sizeof(MjpbrMaterials), (int64_t)sizeof(MjpbrLights),
      (int64_t)sizeof(MjpbrCameras),  (int64_t)sizeof(MjpbrEnvironment), (int64_t)sizeof(MjpbrSettings),
      (int64_t)sizeof(MjpbrOutputs),  (int64_t)sizeof(MjpbrScene),      (int64_t)sizeof(MjpbrAtrousParams),
      (int64_t)sizeof(MjpbrWavefrontConfig), (int64_t)sizeof(MjpbrWavefrontPlan),
  };
  int count = (int)(sizeof(all) / sizeof(all[0]));
  for (int i = 0; i < n && i < count; ++i) sizes[i] = all[i];
  return count;
}
