// Engine kernels (wide nodes, sorted queries).  Placeholder until the wide layout lands.
#include "gq_device.h"
#include "gq_internal.h"
#include "wide.h"

namespace gqa {

hipError_t EngineWorkspace::ensure(size_t need) {
    if (need <= bytes) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&buf, need);
    if (e == hipSuccess) bytes = need;
    return e;
}

void EngineWorkspace::release() {
    if (buf) hipFree(buf);
    buf = nullptr;
    bytes = 0;
}

hipError_t launch_engine(QueryKind, const WideView &, bool, const QueryArgs &, EngineWorkspace &, hipStream_t) {
    set_error("engine method not built yet");
    return hipErrorNotSupported;
}

void build_wide(const HostStructure &, WideHost &) {}

} // namespace gqa
