// reference_caller.cpp -- callers written the way the reference's apps call
// apo, built against include/apo/gpu/algorithm.hpp and the shim of
// apo_gpu_shim.cpp.
//
//   reference_caller <sites.bin> <out.bin> <k> <passes>
//
// sites.bin: n x (x, y, z, r) float32, read into std::vector<apo::Real> as
// the apps hold their sites.  Writes, for two builds in turn, a uint64
// vertex count V, V x (x, y, z, w) float64 and V x (i, j, k, l) uint32:
//   1. apo::gpu::computeFullApolloniusDiagram(sites) through the shim (the
//      precision app's call, app/precision/precision_apo.cpp:17), followed
//      by its closed-edge count E, E x float64 x 4 and E x uint32 x 4;
//   2. AlgorithmGPU<K>{sites, passes}.build() + toHost (the neighbours-per-pass
//      study, app/benchmarks/benchmark_knn.cu:40), K = 1, 4 or 16.
// The scene members of apo::gpu::Algorithm are printed for the second.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "apo/gpu/algorithm.hpp"

namespace {

void writeVertices(std::FILE* out, const std::vector<apo::Real>& xyzw, const std::vector<uint32_t>& ids) {
    const uint64_t v = ids.size() / 4;
    std::fwrite(&v, sizeof(v), 1, out);
    std::fwrite(xyzw.data(), sizeof(double), xyzw.size(), out);
    std::fwrite(ids.data(), sizeof(uint32_t), ids.size(), out);
}

template <uint32_t K>
void knnStudy(const std::vector<apo::Real>& sites, uint32_t passes, std::FILE* out) {
    apo::gpu::AlgorithmGPU<K> algorithm{sites, passes};
    algorithm.build();
    const apo::gpu::VertexDiagram d = algorithm.toHost(true);
    writeVertices(out, d.vertices, d.verticesId);
    std::printf("K=%u passes=%u: %u sites, radii [%g, %g], box [%g %g %g]-[%g %g %g], %zu vertices, knn %.3f ms\n", K,
                algorithm.knnValidationStep, algorithm.m_siteNb, algorithm.m_minRadius, algorithm.m_maxRadius,
                algorithm.m_minX, algorithm.m_minY, algorithm.m_minZ, algorithm.m_maxX, algorithm.m_maxY,
                algorithm.m_maxZ, d.verticesId.size() / 4, algorithm.sample.knnSearch);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s <sites.bin> <out.bin> <k: 1, 4 or 16> <passes>\n", argv[0]);
        return 2;
    }
    std::FILE* in = std::fopen(argv[1], "rb");
    if (!in) {
        std::perror(argv[1]);
        return 1;
    }
    std::vector<apo::Real> sites;
    float buf[4096];
    std::size_t got;
    while ((got = std::fread(buf, sizeof(float), 4096, in)) > 0) sites.insert(sites.end(), buf, buf + got);
    std::fclose(in);
    const uint32_t k = uint32_t(std::atoi(argv[3])), passes = uint32_t(std::atoi(argv[4]));
    std::FILE* out = std::fopen(argv[2], "wb");
    if (!out) {
        std::perror(argv[2]);
        return 1;
    }
    try {
        const apo::gpu::FullDiagram f = apo::gpu::computeFullApolloniusDiagram(sites);
        writeVertices(out, f.vertices, f.verticesId);
        writeVertices(out, f.closedEdgesMin, f.closedEdgesId);
        if (k == 1) knnStudy<1>(sites, passes, out);
        else if (k == 4) knnStudy<4>(sites, passes, out);
        else knnStudy<16>(sites, passes, out);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "apo error: %s\n", e.what());
        std::fclose(out);
        return 1;
    }
    std::fclose(out);
    return 0;
}
