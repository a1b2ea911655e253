// diagram.cpp -- C++ caller of the library through include/apo/gpu/algorithm.hpp.
//
//   diagram <sites.bin> <out.bin> [topology.json]
//
// sites.bin: n x (x, y, z, r) float32.  out.bin: uint64 vertex count V, then
// V x (x, y, z, w) float64, then V x (i, j, k, l) uint32 original site ids,
// then uint64 closed-edge count E, E x (x, y, z, w) float64 and E x
// (i, j, k, 0) uint32.  The call sequence is the reference's
// computeFullApolloniusDiagram (src/gpu/algorithm.cu:17-26); topology.json
// is the document of its topology app (app/topology/topology_apo.cu).
#include <cstdio>
#include <string>
#include <vector>

#include "apo/gpu/algorithm.hpp"

int main(int argc, char** argv) {
    if (argc != 3 && argc != 4) {
        std::fprintf(stderr, "usage: %s <sites.bin> <out.bin> [topology.json]\n", argv[0]);
        return 2;
    }
    std::FILE* in = std::fopen(argv[1], "rb");
    if (!in) {
        std::perror(argv[1]);
        return 1;
    }
    std::vector<float> sites;
    float buf[4096];
    std::size_t got;
    while ((got = std::fread(buf, sizeof(float), 4096, in)) > 0) sites.insert(sites.end(), buf, buf + got);
    std::fclose(in);
    try {
        apo::gpu::AlgorithmGPU<> algo(apo::ConstSpan<float>{sites.data(), sites.size()});
        algo.build();
        const apo::gpu::FullDiagram d = algo.toHostFull(true);
        const uint64_t v = d.vertices.size() / 4;
        const uint64_t e = d.closedEdgesMin.size() / 4;
        std::FILE* out = std::fopen(argv[2], "wb");
        if (!out) {
            std::perror(argv[2]);
            return 1;
        }
        std::fwrite(&v, sizeof(v), 1, out);
        std::fwrite(d.vertices.data(), sizeof(double), d.vertices.size(), out);
        std::fwrite(d.verticesId.data(), sizeof(uint32_t), d.verticesId.size(), out);
        std::fwrite(&e, sizeof(e), 1, out);
        std::fwrite(d.closedEdgesMin.data(), sizeof(double), d.closedEdgesMin.size(), out);
        std::fwrite(d.closedEdgesId.data(), sizeof(uint32_t), d.closedEdgesId.size(), out);
        std::fclose(out);
        if (argc == 4) {
            const std::string json = apo::toJson(algo.toTopology());
            std::FILE* tj = std::fopen(argv[3], "wb");
            if (!tj) {
                std::perror(argv[3]);
                return 1;
            }
            std::fwrite(json.data(), 1, json.size(), tj);
            std::fclose(tj);
        }
        std::printf("%zu sites, %llu vertices, %llu closed edges, %.3f ms (lbvh %.3f, knn %.3f, cells %.3f, validation %.3f)\n",
                    sites.size() / 4, static_cast<unsigned long long>(v),
                    static_cast<unsigned long long>(e), algo.sample.total,
                    algo.sample.bvhConstruction, algo.sample.knnSearch, algo.sample.knnConstruction,
                    algo.sample.vertexValidation + algo.sample.edgeValidation + algo.sample.bisectorValidation);
    } catch (const apo::gpu::Error& e) {
        std::fprintf(stderr, "apo error %d: %s\n", int(e.status), e.what());
        return 1;
    }
    return 0;
}
