// Library-level entry points: ABI version and status strings.
#include "common.h"

#define PAMD_ABI_VERSION 1

extern "C" int pamd_abi_version(void) { return PAMD_ABI_VERSION; }

extern "C" const char* pamd_status_string(int status) {
  if (status == PAMD_OK) return "ok";
  if (status == PAMD_ERR_HIP) return "HIP runtime error";
  if (status == PAMD_ERR_UNSUPPORTED) return "shape or option not supported by parla_amd";
  if (status == PAMD_EVENT_SINGULAR) return "singular triangular factor (exact zero on the diagonal)";
  if (status < 0) return "invalid argument (the status is minus its position)";
  return "numerical event";
}
