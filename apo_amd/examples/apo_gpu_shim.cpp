// apo_gpu_shim.cpp -- the translation unit a maintainer puts in place of
// apo/src/gpu/algorithm.cu (computeApolloniusDiagram and
// computeFullApolloniusDiagram, :6-26) to route the reference's entry
// points to this library.  It sees only the reference's declarations
// (apo/gpu/algorithm.hpp: Real, ConstSpan, VertexDiagram, FullDiagram) and
// the C ABI apo_c.h; here it is compiled against this repository's header
// with APO_GPU_DECLARATIONS_ONLY, which declares the same two functions.
//
// Real (double) sites are rounded to FP32 once; vertices and closed edges
// come back in FP64.  Failures throw std::runtime_error with the library's
// message (the reference prints and continues).
#include <stdexcept>
#include <string>
#include <vector>

#include "apo/gpu/algorithm.hpp"
#include "apo_c.h"

namespace apo::gpu {
namespace {

// AlgorithmGPU<16>(sites, knnValidationStep = 1) + build() (algorithm.inl:1404-1594)
struct Built {
    apo_ctx* ctx = nullptr;
    explicit Built(ConstSpan<Real> sites) {
        std::vector<float> xyzr(sites.begin(), sites.end());
        apo_config cfg;
        apo_default_config(&cfg);
        if (apo_create(&ctx, &cfg) != APO_OK) throw std::runtime_error("apo_create failed");
        if (apo_build(ctx, xyzr.data(), uint32_t(sites.size / 4)) != APO_OK) fail("apo_build");
    }
    ~Built() { apo_destroy(ctx); }
    [[noreturn]] void fail(const char* what) { throw std::runtime_error(std::string(what) + ": " + apo_last_error(ctx)); }

    // toHost (algorithm.inl:1596-1646)
    void vertices(std::vector<Real>& xyzw, std::vector<uint32_t>& ids) {
        uint64_t v = 0;
        if (apo_vertex_count(ctx, &v) != APO_OK) fail("apo_vertex_count");
        std::vector<double> x(4 * v);
        ids.resize(4 * v);
        if (apo_get_vertices(ctx, x.data(), ids.data(), v) != APO_OK) fail("apo_get_vertices");
        xyzw.assign(x.begin(), x.end());
    }
};

}  // namespace

VertexDiagram computeApolloniusDiagram(ConstSpan<Real> sites) {
    Built b(sites);
    VertexDiagram d;
    b.vertices(d.vertices, d.verticesId);
    return d;
}

FullDiagram computeFullApolloniusDiagram(ConstSpan<Real> sites) {
    Built b(sites);
    FullDiagram d;
    b.vertices(d.vertices, d.verticesId);
    // toHostFull's closed trisector edges (algorithm.inl:1648-1730)
    uint64_t e = 0;
    if (apo_closed_edge_count(b.ctx, &e) != APO_OK) b.fail("apo_closed_edge_count");
    std::vector<double> x(4 * e);
    d.closedEdgesId.resize(4 * e);
    if (e && apo_get_closed_edges(b.ctx, x.data(), d.closedEdgesId.data(), e) != APO_OK) b.fail("apo_get_closed_edges");
    d.closedEdgesMin.assign(x.begin(), x.end());
    return d;
}

}  // namespace apo::gpu
