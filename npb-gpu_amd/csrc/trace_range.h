// roctx ranges around the library's top-level calls (HIP sources only).
#ifndef NPBCG_TRACE_RANGE_H
#define NPBCG_TRACE_RANGE_H

#include <rocprofiler-sdk-roctx/roctx.h>

namespace npbcg_detail {

// A roctx range named after a library call, open for the scope: a marker
// trace (rocprofv3 --marker-trace) then attributes each kernel to the call
// that launched it.
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

}  // namespace npbcg_detail

#endif
