// samples.cpp -- seeded protein-like molecules (host C++, no GPU).
//
// The synthetic stand-in for the reference's protein inputs
// (app/common/samples.cpp:87-127 loads MMTF files with chemfiles, which are
// not available here): backbone chains are random walks of 3.8 A steps
// with bond angle uniform in [80, 150] degrees and torsion uniform in
// [0, 360); each residue carries 1-4 side atoms at 1.5 A in random
// directions; an atom closer than 1.2 A to an existing one is redrawn (up to
// 24 times, then the chain ends), and a step leaving the cube also ends the
// chain.  Chains start at random free points until n atoms are placed in a
// cube of side (n / density)^(1/3).  Elements C / N / O / S are drawn with
// frequencies .62 / .17 / .19 / .02 and get radii 1.70 / 1.55 / 1.52 / 1.80 A.
//
// Randomness: PCG32 (O'Neill's pcg32, XSH-RR output, 64-bit LCG state)
// seeded with (seed, stream); uniform doubles take 53 bits of two outputs,
// random directions are uniform on the sphere (z uniform in [-1, 1],
// azimuth uniform).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct Pcg32 {
    uint64_t state = 0, inc = 1;
    Pcg32(uint64_t seed, uint64_t stream) {
        inc = (stream << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t x = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (x >> rot) | (x << ((32u - rot) & 31u));
    }
    // uniform in [0, 1) with 53 random bits
    double unit() {
        const uint64_t hi = next() >> 5, lo = next() >> 6;
        return double((hi << 26) | lo) * (1.0 / 9007199254740992.0);
    }
    double uniform(double a, double b) { return a + (b - a) * unit(); }
    // integer in [a, b)
    int range(int a, int b) { return a + int(unit() * double(b - a)); }
};

struct V3 {
    double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

V3 direction(Pcg32& g) {
    const double z = g.uniform(-1.0, 1.0), t = g.uniform(0.0, 2.0 * M_PI);
    const double r = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    return {r * std::cos(t), r * std::sin(t), z};
}

// Uniform grid of 1.2 A cells with a linked list of atoms per cell.
struct Grid {
    double side, cell;
    int dims;
    std::vector<int32_t> head, next;
    std::vector<V3> pos;
    Grid(double s, double c, size_t n) : side(s), cell(c) {
        dims = std::max(1, int(s / c));
        head.assign(size_t(dims) * dims * dims, -1);
        next.reserve(n);
        pos.reserve(n);
    }
    int key(double v) const { return std::min(int(v / cell), dims - 1); }
    bool free(V3 p) const {
        if (!(p.x >= 0.0 && p.x < side && p.y >= 0.0 && p.y < side && p.z >= 0.0 && p.z < side)) return false;
        const int kx = key(p.x), ky = key(p.y), kz = key(p.z);
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    const int x = kx + dx, y = ky + dy, z = kz + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= dims || y >= dims || z >= dims) continue;
                    for (int32_t q = head[(size_t(x) * dims + y) * dims + z]; q >= 0; q = next[size_t(q)]) {
                        const V3 d = p - pos[size_t(q)];
                        if (dot(d, d) < cell * cell) return false;
                    }
                }
        return true;
    }
    void add(V3 p) {
        const size_t c = (size_t(key(p.x)) * dims + key(p.y)) * dims + key(p.z);
        next.push_back(head[c]);
        head[c] = int32_t(pos.size());
        pos.push_back(p);
    }
};

}  // namespace

extern "C" {

// n protein-like atoms as n x (x, y, z, r) FP32 into out.  Returns 0, or -1
// for invalid arguments.
int apo_samples_protein(uint32_t n, uint64_t seed, uint64_t stream, double density, float* out) {
    if (!out || n == 0 || !(density > 0.0)) return -1;
    Pcg32 g(seed, stream);
    const double side = std::cbrt(double(n) / density);
    Grid grid(side, 1.2, n);
    while (grid.pos.size() < n) {
        const V3 p0{g.uniform(0.0, side), g.uniform(0.0, side), g.uniform(0.0, side)};
        if (!grid.free(p0)) continue;
        grid.add(p0);
        V3 prev = p0, prev2{0, 0, 0};
        bool second = false;
        const int residues = g.range(50, 400);
        for (int r = 0; r < residues && grid.pos.size() < n; r++) {
            bool placed = false;
            V3 q{0, 0, 0};
            for (int t = 0; t < 24 && !placed; t++) {
                if (!second) {
                    q = prev + 3.8 * direction(g);
                } else {
                    // next backbone atom from bond angle and torsion (NeRF)
                    const double theta = g.uniform(80.0, 150.0) * (M_PI / 180.0);
                    const double phi = g.uniform(0.0, 2.0 * M_PI);
                    V3 b = prev - prev2;
                    b = (1.0 / std::sqrt(dot(b, b))) * b;
                    V3 nrm = cross(b, direction(g));
                    const double nn = std::sqrt(dot(nrm, nrm));
                    if (nn < 1e-6) continue;
                    nrm = (1.0 / nn) * nrm;
                    const V3 m = cross(nrm, b);
                    const double d0 = -std::cos(theta), d1 = std::sin(theta) * std::cos(phi),
                                 d2 = std::sin(theta) * std::sin(phi);
                    q = prev + 3.8 * (d0 * b + d1 * m + d2 * nrm);
                }
                placed = grid.free(q);
            }
            if (!placed) break;
            grid.add(q);
            prev2 = prev;
            prev = q;
            second = true;
            const int sides = g.range(1, 5);
            for (int s = 0; s < sides && grid.pos.size() < n; s++)
                for (int t = 0; t < 24; t++) {
                    const V3 a = q + 1.5 * direction(g);
                    if (grid.free(a)) {
                        grid.add(a);
                        break;
                    }
                }
        }
    }
    static const double radii[4] = {1.70, 1.55, 1.52, 1.80};
    static const double cumulative[4] = {0.62, 0.79, 0.98, 1.0};
    for (uint32_t i = 0; i < n; i++) {
        const double u = g.unit();
        int e = 0;
        while (e < 3 && u >= cumulative[e]) e++;
        const V3 p = grid.pos[i];
        out[4 * size_t(i) + 0] = float(p.x);
        out[4 * size_t(i) + 1] = float(p.y);
        out[4 * size_t(i) + 2] = float(p.z);
        out[4 * size_t(i) + 3] = float(radii[e]);
    }
    return 0;
}

}  // extern "C"
