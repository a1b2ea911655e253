// apo/gpu/algorithm.hpp -- C++17 host interface over the C ABI (apo_c.h).
//
// Restores the construction API of the reference, apo/include/apo/gpu/
// algorithm.hpp:10-59 and algorithm.cuh:12-46, so a caller written against
// the reference compiles unchanged with this include path:
//   apo::Real (double, or float with APO_REAL_SIZE=4: core/math.hpp:27-31),
//   apo::Span / apo::ConstSpan (core/type.hpp:8-40), apo::joggle
//   (core/utils.hpp:9, src/core/utils.cpp:14-25);
//   VertexDiagram, FullDiagram, Topology, the Algorithm base with its scene
//   members (src/gpu/algorithm.cpp:5-25), AlgorithmGPU<K>(sites,
//   knnValidationStep = 1) with build, toHost, toHostFull, toTopology and the
//   stage-time Sample, computeApolloniusDiagram and
//   computeFullApolloniusDiagram;
// and the topology document of the reference's topology app (toCsv,
// app/topology/topology.cpp:11-57) as apo::toJson.
//
// Sites are (x, y, z, r) quadruples, 4n values.  The engine computes on FP32
// sites: Real (double) input is rounded to FP32 once, at build(); FP32
// input (the ConstSpan<float> overloads) is used as given.  Vertices come
// back as Real (x, y, z, w) from the engine's FP64 results, w the
// empty-sphere radius, with their four tangent sites as original indices.
//
// Differences from the reference: failures throw apo::gpu::Error (the
// reference prints and continues, setup.inl:8-20); the device state is
// reached through the introspection calls of apo_c.h instead of public
// DeviceBuffer members; Sample::initialization is 0 (the octahedra are
// written by the first cell pass, whose time is knnConstruction).
//
// With APO_GPU_DECLARATIONS_ONLY defined, computeApolloniusDiagram and
// computeFullApolloniusDiagram are declared, not defined, as in the
// reference's header: a translation unit of the caller's own then defines
// them (apo_amd/examples/apo_gpu_shim.cpp).
#ifndef APO_GPU_ALGORITHM_HPP
#define APO_GPU_ALGORITHM_HPP

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "apo_c.h"

namespace apo {

#if defined(APO_REAL_SIZE) && APO_REAL_SIZE == 4
using Real = float;
#else
using Real = double;
#endif

template <class Type>
struct Span {
    Type* ptr = nullptr;
    std::size_t size = 0;

    Span() = default;
    Span(Type* pptr, std::size_t psize) : ptr(pptr), size(psize) {}
    Span(std::vector<Type>& data) : ptr(data.data()), size(data.size()) {}
    template <std::size_t N>
    Span(std::array<Type, N>& data) : ptr(data.data()), size(N) {}

    Type& operator[](std::size_t i) const { return ptr[i]; }
    Type* begin() const { return ptr; }
    Type* end() const { return ptr + size; }
};

template <class Type>
struct ConstSpan {
    const Type* ptr = nullptr;
    std::size_t size = 0;

    ConstSpan() = default;
    ConstSpan(const Type* pptr, std::size_t psize) : ptr(pptr), size(psize) {}
    ConstSpan(const std::vector<Type>& data) : ptr(data.data()), size(data.size()) {}
    template <std::size_t N>
    ConstSpan(const std::array<Type, N>& data) : ptr(data.data()), size(N) {}
    ConstSpan(const Span<Type> data) : ptr(data.ptr), size(data.size) {}

    const Type& operator[](std::size_t i) const { return ptr[i]; }
    const Type* begin() const { return ptr; }
    const Type* end() const { return ptr + size; }
};

// apo::joggle (src/core/utils.cpp:14-25): every value v[4i + j] plus
// pcg(4i + j) * 2 range - range, general position for the construction.
inline void joggle(Span<double> sites, double range = 1e-4) { apo_joggle(sites.ptr, sites.size, range); }
// FP32 sites: the same noise in FP64, each value rounded to FP32 once.
inline void joggle(Span<float> sites, double range = 1e-4) { apo_joggle_f32(sites.ptr, sites.size, range); }

}  // namespace apo

namespace apo::gpu {

class Error : public std::runtime_error {
  public:
    Error(apo_status status, const std::string& what) : std::runtime_error(what), status(status) {}
    apo_status status;
};

struct VertexDiagram {
    std::vector<Real> vertices;            // Real4 {x, y, z, w}
    std::vector<uint32_t> verticesId;      // {i, j, k, l} per vertex, original site indices
    std::vector<uint32_t> invalidationId;  // never filled (as in the reference, algorithm.hpp:14)
};

// FullDiagram (algorithm.hpp:18-26): the vertices plus each closed trisector
// edge as its extreme point of smaller empty-sphere radius.
struct FullDiagram {
    std::vector<Real> vertices;            // Real4 {x, y, z, w}
    std::vector<uint32_t> verticesId;      // {i, j, k, l} per vertex
    std::vector<uint32_t> invalidationId;  // never filled
    std::vector<uint32_t> closedEdgesId;   // {i, j, k, 0} per closed edge
    std::vector<Real> closedEdgesMin;      // Real4 {x, y, z, w} per closed edge
};

// Topology (algorithm.hpp:28-33), original site ids.
struct Topology {
    std::vector<uint32_t> quadrisectors;  // uint4
    std::vector<uint32_t> trisectors;     // uint3
    std::vector<uint32_t> bisectors;      // uint2
};

// Algorithm (algorithm.hpp:35-56): the sites and their scene members, each
// computed as src/gpu/algorithm.cpp:5-25 does (extents x - r ... x + r,
// smallest and largest radius, in Real).
class Algorithm {
  public:
    explicit Algorithm(ConstSpan<Real> sites) : m_sites(sites), m_siteNb(uint32_t(sites.size / 4)) {
        for (std::size_t v = 0; v < m_siteNb; v++) extend(sites[v * 4 + 0], sites[v * 4 + 1], sites[v * 4 + 2], sites[v * 4 + 3]);
    }
    virtual ~Algorithm() = default;

    virtual void build() = 0;
    virtual VertexDiagram toHost(bool validate) = 0;
    virtual FullDiagram toHostFull(bool validate) = 0;
    virtual Topology toTopology() = 0;

    ConstSpan<Real> m_sites;     // the caller's Real sites (empty for FP32 input)
    ConstSpan<float> m_fsites;   // the caller's FP32 sites (empty for Real input)
    uint32_t m_siteNb = 0;
    Real m_minRadius = std::numeric_limits<Real>::max();
    Real m_maxRadius = -std::numeric_limits<Real>::max();
    Real m_minX = std::numeric_limits<Real>::max();
    Real m_minY = std::numeric_limits<Real>::max();
    Real m_minZ = std::numeric_limits<Real>::max();
    Real m_maxX = -std::numeric_limits<Real>::max();
    Real m_maxY = -std::numeric_limits<Real>::max();
    Real m_maxZ = -std::numeric_limits<Real>::max();

  protected:
    // FP32 sites (no counterpart in the reference, whose input is Real)
    explicit Algorithm(ConstSpan<float> sites, int) : m_fsites(sites), m_siteNb(uint32_t(sites.size / 4)) {
        for (std::size_t v = 0; v < m_siteNb; v++)
            extend(Real(sites[v * 4 + 0]), Real(sites[v * 4 + 1]), Real(sites[v * 4 + 2]), Real(sites[v * 4 + 3]));
    }

  private:
    void extend(Real x, Real y, Real z, Real r) {
        m_minX = std::min(m_minX, x - r);
        m_minY = std::min(m_minY, y - r);
        m_minZ = std::min(m_minZ, z - r);
        m_maxX = std::max(m_maxX, x + r);
        m_maxY = std::max(m_maxY, y + r);
        m_maxZ = std::max(m_maxZ, z + r);
        m_maxRadius = std::max(m_maxRadius, r);
        m_minRadius = std::min(m_minRadius, r);
    }
};

// AlgorithmGPU<K> (algorithm.cuh:12-46): K neighbours per kNN pass (1..64),
// knnValidationStep kNN + cell passes (algorithm.inl:1535-1549).
template <uint32_t K = 16>
class AlgorithmGPU : public Algorithm {
    static_assert(K >= 1 && K <= 64, "AlgorithmGPU<K>: K neighbours per pass, 1..64");

  public:
    AlgorithmGPU(ConstSpan<Real> sites, uint32_t knnValidationStep = 1)
        : Algorithm(sites), knnValidationStep(knnValidationStep) {
        create(sites.size);
    }
    template <class T = Real, class = std::enable_if_t<!std::is_same_v<T, float>>>
    AlgorithmGPU(ConstSpan<float> sites, uint32_t knnValidationStep = 1)
        : Algorithm(sites, 0), knnValidationStep(knnValidationStep) {
        create(sites.size);
    }
    AlgorithmGPU(const AlgorithmGPU&) = delete;
    AlgorithmGPU& operator=(const AlgorithmGPU&) = delete;
    ~AlgorithmGPU() override { apo_destroy(m_ctx); }

    // Stage split of one build in milliseconds (AlgorithmGPU::Sample,
    // algorithm.cuh:18-27), plus the output stage and the total.
    struct Sample {
        double initialization = 0, bvhConstruction = 0, knnSearch = 0, knnConstruction = 0, vertexValidation = 0,
               edgeValidation = 0, bisectorValidation = 0, output = 0, total = 0;
    };

    void build() override {
        const float* xyzr = m_fsites.ptr;
        if (!xyzr) {
            if constexpr (std::is_same_v<Real, float>) {
                xyzr = m_sites.ptr;
            } else {
                m_input.assign(m_sites.begin(), m_sites.end());  // each value rounded to FP32 once
                xyzr = m_input.data();
            }
        }
        check(apo_build(m_ctx, xyzr, m_siteNb), "apo_build");
        apo_stats s;
        check(apo_get_stats(m_ctx, &s), "apo_get_stats");
        sample = Sample{};
        sample.bvhConstruction = s.ms_lbvh;
        sample.knnSearch = s.ms_knn;
        sample.knnConstruction = s.ms_cells;
        sample.vertexValidation = s.ms_vertex_validation_kernel;
        sample.edgeValidation = s.ms_edge_validation_kernel;
        sample.bisectorValidation = s.ms_bisector;
        sample.output = s.ms_output;
        sample.total = s.ms_total;
    }

    // The flag is ignored, as in the reference (algorithm.inl:1596).
    VertexDiagram toHost(bool /*validate*/ = true) override {
        uint64_t count = 0;
        check(apo_vertex_count(m_ctx, &count), "apo_vertex_count");
        VertexDiagram d;
        d.verticesId.resize(4 * count);
        if constexpr (std::is_same_v<Real, double>) {
            d.vertices.resize(4 * count);
            check(apo_get_vertices(m_ctx, d.vertices.data(), d.verticesId.data(), count), "apo_get_vertices");
        } else {
            std::vector<double> xyzw(4 * count);
            check(apo_get_vertices(m_ctx, xyzw.data(), d.verticesId.data(), count), "apo_get_vertices");
            d.vertices.assign(xyzw.begin(), xyzw.end());
        }
        return d;
    }

    // toHostFull (algorithm.inl:1648-1730); the flag is ignored.
    FullDiagram toHostFull(bool validate = true) override {
        VertexDiagram v = toHost(validate);
        FullDiagram d;
        d.vertices = std::move(v.vertices);
        d.verticesId = std::move(v.verticesId);
        uint64_t count = 0;
        check(apo_closed_edge_count(m_ctx, &count), "apo_closed_edge_count");
        std::vector<double> xyzw(4 * count);
        d.closedEdgesId.resize(4 * count);
        if (count)
            check(apo_get_closed_edges(m_ctx, xyzw.data(), d.closedEdgesId.data(), count), "apo_get_closed_edges");
        d.closedEdgesMin.assign(xyzw.begin(), xyzw.end());
        return d;
    }

    // toTopology (algorithm.inl:1732-1821)
    Topology toTopology() override {
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

    uint32_t knnValidationStep;
    Sample sample;

  private:
    void create(std::size_t values) {
        if (values % 4 != 0) throw Error(APO_ERR_INVALID, "site array length is not a multiple of 4");
        apo_config cfg;
        apo_default_config(&cfg);
        cfg.k = K;
        cfg.knn_passes = knnValidationStep;
        check(apo_create(&m_ctx, &cfg), "apo_create");
    }
    void check(apo_status st, const char* what) {
        if (st != APO_OK)
            throw Error(st, std::string(what) + ": " + (m_ctx ? apo_last_error(m_ctx) : "invalid configuration"));
    }

    std::vector<float> m_input;  // FP32 copy of Real sites
    apo_ctx* m_ctx = nullptr;
};

// computeApolloniusDiagram / computeFullApolloniusDiagram
// (algorithm.hpp:58-59, src/gpu/algorithm.cu:6-26): AlgorithmGPU<16> with
// one kNN pass, build, toHost / toHostFull.
VertexDiagram computeApolloniusDiagram(ConstSpan<Real> sites);
FullDiagram computeFullApolloniusDiagram(ConstSpan<Real> sites);

#ifndef APO_GPU_DECLARATIONS_ONLY
inline VertexDiagram computeApolloniusDiagram(ConstSpan<Real> sites) {
    AlgorithmGPU<> algo(sites);
    algo.build();
    return algo.toHost(true);
}

inline FullDiagram computeFullApolloniusDiagram(ConstSpan<Real> sites) {
    AlgorithmGPU<> algo(sites);
    algo.build();
    return algo.toHostFull(true);
}
#endif

// FP32 sites, no conversion (no counterpart in the reference)
template <class T = Real, class = std::enable_if_t<!std::is_same_v<T, float>>>
inline VertexDiagram computeApolloniusDiagram(ConstSpan<float> sites) {
    AlgorithmGPU<> algo(sites);
    algo.build();
    return algo.toHost(true);
}

template <class T = Real, class = std::enable_if_t<!std::is_same_v<T, float>>>
inline FullDiagram computeFullApolloniusDiagram(ConstSpan<float> sites) {
    AlgorithmGPU<> algo(sites);
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
