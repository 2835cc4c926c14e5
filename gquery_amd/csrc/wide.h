// Engine layout: wide (collapsed) BVH / SNCH nodes in 64-byte records, fp32 and quantized
// variants, primitives and silhouettes in leaf order as float4 records; plus the per-scene
// workspace the engine launcher keeps between calls.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "builder.h"
#include "gq_internal.h"

namespace gqa {

struct WideHost {
    int width = 4;
    size_t n_nodes = 0, n_snch_nodes = 0;
    float scene_lo[3] = {0, 0, 0}, scene_hi[3] = {0, 0, 0};
    std::vector<uint8_t> nodes, qnodes, snodes, qsnodes;  // raw 64-byte records
    std::vector<float4> prims;
    std::vector<int32_t> prim_ids;
    std::vector<float4> sils;
    std::vector<int32_t> sil_ids;
};

void build_wide(const HostStructure &h, WideHost &out);

// Device scratch owned by a scene: query keys / permutation for the Morton sort.  Grown on
// demand (hipMalloc) and reused by later calls on the same stream order.
struct EngineWorkspace {
    void *buf = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need);
    void release();
};

hipError_t launch_engine(QueryKind kind, const WideView &w, bool quantized, const QueryArgs &q,
                         EngineWorkspace &ws, hipStream_t st);

} // namespace gqa
