// A roctx range around one C-ABI call, named after the entry point, so that
// profiler traces (rocprofv3 --marker-trace) show each call with the kernels
// it launched.  Without a profiler attached the push and pop return at once.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace csx {

class ApiRange {
 public:
  explicit ApiRange(const char* name) { roctxRangePush(name); }
  ~ApiRange() { roctxRangePop(); }
  ApiRange(const ApiRange&) = delete;
  ApiRange& operator=(const ApiRange&) = delete;
};

}  // namespace csx

#define CSX_API_RANGE() ::csx::ApiRange csx_api_range_(__func__)
