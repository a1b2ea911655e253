// apo/gpu/algorithm.hpp -- C++17 host interface over the C ABI (apo_c.h).
//
// Restores the construction API of the reference, apo/include/apo/gpu/
// algorithm.hpp:10-59 and algorithm.cuh:12-46: VertexDiagram, FullDiagram,
// Topology, AlgorithmGPU (build, toHost, toHostFull, toTopology, the
// stage-time Sample), computeApolloniusDiagram and
// computeFullApolloniusDiagram; and the topology document of the reference's
// topology app (toCsv, app/topology/topology.cpp:11-57) as apo::toJson.
// Sites are (x, y, z, r) FP32 quadruples, 4n floats; vertices come back as
// FP64 (x, y, z, w) with w the empty-sphere radius, and their four tangent
// sites as original input indices.
//
// Differences from the reference: failures throw apo::gpu::Error (the
// reference prints and continues, setup.inl:8-20); the device state is
// reached through the introspection calls of apo_c.h instead of public
// DeviceBuffer members.
#ifndef APO_GPU_ALGORITHM_HPP
#define APO_GPU_ALGORITHM_HPP

#include <cstddef>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "apo_c.h"

namespace apo::gpu {

class Error : public std::runtime_error {
  public:
    Error(apo_status status, const std::string& what) : std::runtime_error(what), status(status) {}
    apo_status status;
};

// Non-owning view of 4n site scalars (the reference's ConstSpan<Real>, type.hpp:25-41).
struct ConstSpan {
    const float* ptr = nullptr;
    std::size_t size = 0;  // number of floats, 4 per site
};

struct VertexDiagram {
    std::vector<double> vertices;          // {x, y, z, w} per vertex
    std::vector<uint32_t> verticesId;      // {i, j, k, l} per vertex, original site indices
    std::vector<uint32_t> invalidationId;  // never filled (as in the reference, algorithm.hpp:14)
};

// FullDiagram (algorithm.hpp:18-26): the vertices plus each closed trisector
// edge as its extreme point of smaller empty-sphere radius.
struct FullDiagram {
    std::vector<double> vertices;          // {x, y, z, w} per vertex
    std::vector<uint32_t> verticesId;      // {i, j, k, l} per vertex
    std::vector<uint32_t> invalidationId;  // never filled
    std::vector<uint32_t> closedEdgesId;   // {i, j, k, 0} per closed edge
    std::vector<double> closedEdgesMin;    // {x, y, z, w} per closed edge
};

// Topology (algorithm.hpp:28-33), original site ids.
struct Topology {
    std::vector<uint32_t> quadrisectors;  // uint4
    std::vector<uint32_t> trisectors;     // uint3
    std::vector<uint32_t> bisectors;      // uint2
};

class AlgorithmGPU {
  public:
    // Stage split of one build in milliseconds (AlgorithmGPU::Sample, algorithm.cuh:18-27).
    struct Sample {
        double bvhConstruction = 0, knnSearch = 0, knnConstruction = 0, vertexEdgeValidation = 0,
               bisectorValidation = 0, output = 0, total = 0;
    };

    explicit AlgorithmGPU(ConstSpan sites, uint32_t k = 16) : m_sites(sites) {
        if (sites.size % 4 != 0) throw Error(APO_ERR_INVALID, "site array length is not a multiple of 4");
        apo_config cfg;
        apo_default_config(&cfg);
        cfg.k = k;
        check(apo_create(&m_ctx, &cfg), "apo_create");
    }
    AlgorithmGPU(const AlgorithmGPU&) = delete;
    AlgorithmGPU& operator=(const AlgorithmGPU&) = delete;
    ~AlgorithmGPU() { apo_destroy(m_ctx); }

    void build() {
        check(apo_build(m_ctx, m_sites.ptr, uint32_t(m_sites.size / 4)), "apo_build");
        apo_stats s;
        check(apo_get_stats(m_ctx, &s), "apo_get_stats");
        sample.bvhConstruction = s.ms_lbvh;
        sample.knnSearch = s.ms_knn;
        sample.knnConstruction = s.ms_cells;
        sample.vertexEdgeValidation = s.ms_validation;
        sample.bisectorValidation = s.ms_bisector;
        sample.output = s.ms_output;
        sample.total = s.ms_total;
    }

    // The flag is ignored, as in the reference (algorithm.inl:1596).
    VertexDiagram toHost(bool /*validate*/ = true) {
        uint64_t count = 0;
        check(apo_vertex_count(m_ctx, &count), "apo_vertex_count");
        VertexDiagram d;
        d.vertices.resize(4 * count);
        d.verticesId.resize(4 * count);
        check(apo_get_vertices(m_ctx, d.vertices.data(), d.verticesId.data(), count), "apo_get_vertices");
        return d;
    }

    // toHostFull (algorithm.inl:1648-1730); the flag is ignored.
    FullDiagram toHostFull(bool validate = true) {
        VertexDiagram v = toHost(validate);
        FullDiagram d;
        d.vertices = std::move(v.vertices);
        d.verticesId = std::move(v.verticesId);
        uint64_t count = 0;
        check(apo_closed_edge_count(m_ctx, &count), "apo_closed_edge_count");
        d.closedEdgesMin.resize(4 * count);
        d.closedEdgesId.resize(4 * count);
        if (count)
            check(apo_get_closed_edges(m_ctx, d.closedEdgesMin.data(), d.closedEdgesId.data(), count),
                  "apo_get_closed_edges");
        return d;
    }

    // toTopology (algorithm.inl:1732-1821)
    Topology toTopology() {
        uint64_t c[3] = {0, 0, 0};
        check(apo_get_topology_sets(m_ctx, nullptr, 0, nullptr, 0, nullptr, 0, c), "apo_get_topology_sets");
        Topology t;
        t.bisectors.resize(2 * c[0] + 2);
        t.trisectors.resize(3 * c[1] + 3);
        t.quadrisectors.resize(4 * c[2] + 4);
        check(apo_get_topology_sets(m_ctx, t.bisectors.data(), c[0], t.trisectors.data(), c[1],
                                    t.quadrisectors.data(), c[2], c),
              "apo_get_topology_sets");
        t.bisectors.resize(2 * c[0]);
        t.trisectors.resize(3 * c[1]);
        t.quadrisectors.resize(4 * c[2]);
        return t;
    }

    apo_stats stats() {
        apo_stats s;
        check(apo_get_stats(m_ctx, &s), "apo_get_stats");
        return s;
    }

    apo_ctx* handle() { return m_ctx; }

    Sample sample;

  private:
    void check(apo_status st, const char* what) {
        if (st != APO_OK)
            throw Error(st, std::string(what) + ": " + (m_ctx ? apo_last_error(m_ctx) : "no context"));
    }

    ConstSpan m_sites;
    apo_ctx* m_ctx = nullptr;
};

// computeApolloniusDiagram (src/gpu/algorithm.cu:6-15)
inline VertexDiagram computeApolloniusDiagram(ConstSpan sites) {
    AlgorithmGPU algo(sites);
    algo.build();
    return algo.toHost(true);
}

// computeFullApolloniusDiagram (src/gpu/algorithm.cu:17-26)
inline FullDiagram computeFullApolloniusDiagram(ConstSpan sites) {
    AlgorithmGPU algo(sites);
    algo.build();
    return algo.toHostFull(true);
}

}  // namespace apo::gpu

namespace apo {

// The topology document the reference's topology app writes (toCsv,
// app/topology/topology.cpp:11-57): bisectors, trisectors, quadrisectors.
inline std::string toJson(const gpu::Topology& t) {
    std::ostringstream o;
    auto section = [&o](const char* name, const std::vector<uint32_t>& v, std::size_t width, bool last) {
        o << "\t\"" << name << "\": [\n";
        const std::size_t m = v.size() / width;
        for (std::size_t r = 0; r < m; r++) {
            o << "\t\t[ ";
            for (std::size_t c = 0; c < width; c++) o << v[r * width + c] << (c + 1 < width ? ", " : " ]");
            o << (r + 1 == m ? "" : ",") << "\n";
        }
        o << (last ? "\t]\n" : "\t],\n");
    };
    o << "{\n";
    section("bisectors", t.bisectors, 2, false);
    section("trisectors", t.trisectors, 3, false);
    section("quadrisectors", t.quadrisectors, 4, true);
    o << "}\n";
    return o.str();
}

}  // namespace apo

#endif  // APO_GPU_ALGORITHM_HPP
