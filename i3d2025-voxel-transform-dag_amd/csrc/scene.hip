// Seeded synthetic scenes (host code): shells of spheres and axis-aligned boxes, a value-noise
// heightfield, and a city of boxes with window voids. Output is a dense bit grid in the
// VoxelGrid layout (bit i = x + R*y + R*R*z, voxel_grid.h:70-73). The configurations that use
// each scene kind are described in configs/*.json.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "pipeline.hpp"

namespace tdag {

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // 24-bit uniform in [0, 1).
    double uniform() { return (double)(next() >> 40) * (1.0 / 16777216.0); }
};

uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Sphere {
    double c[3], r;
};
struct Box {
    double lo[3], e[3];
};

inline void setBit(uint8_t* bits, uint64_t i) { bits[i >> 3] |= (uint8_t)(1u << (i & 7)); }

template <typename F>
void parallelSlabs(uint32_t R, int threads, F&& f)
{
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t]() {
            for (uint32_t z = (uint32_t)t; z < R; z += (uint32_t)threads)
                f(z);
        });
    for (auto& th : pool)
        th.join();
}

double sphereSdf(const Sphere& s, double px, double py, double pz)
{
    const double dx = px - s.c[0], dy = py - s.c[1], dz = pz - s.c[2];
    return std::sqrt(dx * dx + dy * dy + dz * dz) - s.r;
}

double boxSdf(const Box& b, double px, double py, double pz)
{
    const double p[3] = { px, py, pz };
    double q[3];
    for (int a = 0; a < 3; ++a) {
        const double h = 0.5 * b.e[a];
        q[a] = std::fabs(p[a] - (b.lo[a] + h)) - h;
    }
    const double ox = std::max(q[0], 0.0), oy = std::max(q[1], 0.0), oz = std::max(q[2], 0.0);
    const double outside = std::sqrt(ox * ox + oy * oy + oz * oz);
    const double inside = std::min(std::max(q[0], std::max(q[1], q[2])), 0.0);
    return outside + inside;
}

// Value noise: lattice values from a hash of (seed, octave, ix, iz), smoothstep-bilinear.
double latticeValue(uint64_t seed, uint32_t octave, int64_t ix, int64_t iz)
{
    const uint64_t h = mix64(seed ^ mix64((uint64_t)octave * 0x100000001B3ull ^ mix64((uint64_t)ix * 0x9E3779B97F4A7C15ull ^ (uint64_t)iz)));
    return (double)(h >> 40) * (1.0 / 16777216.0);
}

double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

double valueNoise(uint64_t seed, uint32_t octave, double x, double z)
{
    const double fx = std::floor(x), fz = std::floor(z);
    const int64_t ix = (int64_t)fx, iz = (int64_t)fz;
    const double tx = smooth(x - fx), tz = smooth(z - fz);
    const double v00 = latticeValue(seed, octave, ix, iz), v10 = latticeValue(seed, octave, ix + 1, iz);
    const double v01 = latticeValue(seed, octave, ix, iz + 1), v11 = latticeValue(seed, octave, ix + 1, iz + 1);
    const double a = v00 + (v10 - v00) * tx;
    const double b = v01 + (v11 - v01) * tx;
    return a + (b - a) * tz;
}

// Fractal heightfield column heights: `octaves` octaves, persistence 0.5, base lattice cell
// R/4 voxels, normalised to [0, 1) and scaled to `heightRange` voxels.
std::vector<int32_t> heightfieldTops(uint32_t R, uint64_t seed, uint32_t octaves, double heightRange, int threads)
{
    std::vector<int32_t> top((size_t)R * R);
    const double baseCell = std::max(1.0, R / 4.0);
    double ampSum = 0.0;
    for (uint32_t o = 0; o < octaves; ++o)
        ampSum += std::pow(0.5, (double)o);
    parallelSlabs(R, threads, [&](uint32_t z) {
        for (uint32_t x = 0; x < R; ++x) {
            double v = 0.0, amp = 1.0, cell = baseCell;
            for (uint32_t o = 0; o < octaves; ++o) {
                v += amp * valueNoise(seed, o, (x + 0.5) / cell, (z + 0.5) / cell);
                amp *= 0.5;
                cell *= 0.5;
            }
            const double h = heightRange * (v / ampSum);
            top[(size_t)x + (size_t)R * z] = (int32_t)std::min<double>(R, std::max(0.0, std::floor(h)));
        }
    });
    return top;
}

// Surface shell of the solid {y < top(x, z)}: a solid voxel with a non-solid 6-neighbour
// (neighbours outside the grid count as solid).
void rasterizeTops(uint8_t* bits, uint32_t R, const std::vector<int32_t>& top, int threads)
{
    parallelSlabs(R, threads, [&](uint32_t z) {
        for (uint32_t x = 0; x < R; ++x) {
            const int32_t t = top[(size_t)x + (size_t)R * z];
            if (t <= 0)
                continue;
            int32_t lo = t - 1;
            const int dx[4] = { -1, 1, 0, 0 }, dz[4] = { 0, 0, -1, 1 };
            for (int k = 0; k < 4; ++k) {
                const int64_t nx = (int64_t)x + dx[k], nz = (int64_t)z + dz[k];
                if (nx < 0 || nz < 0 || nx >= R || nz >= R)
                    continue;
                lo = std::min(lo, top[(size_t)nx + (size_t)R * nz]);
            }
            lo = std::max(lo, 0);
            for (int32_t y = lo; y < t; ++y)
                setBit(bits, (uint64_t)x + (uint64_t)R * y + (uint64_t)R * R * z);
        }
    });
}

void rasterizeShapes(uint8_t* bits, uint32_t R, const std::vector<Sphere>& spheres, const std::vector<Box>& boxes, int threads)
{
    parallelSlabs(R, threads, [&](uint32_t z) {
        const double pz = z + 0.5;
        for (const Sphere& s : spheres) {
            if (pz < s.c[2] - s.r - 1.0 || pz > s.c[2] + s.r + 1.0)
                continue;
            const int64_t y0 = std::max<int64_t>(0, (int64_t)std::floor(s.c[1] - s.r - 1.0));
            const int64_t y1 = std::min<int64_t>(R - 1, (int64_t)std::ceil(s.c[1] + s.r + 1.0));
            const int64_t x0 = std::max<int64_t>(0, (int64_t)std::floor(s.c[0] - s.r - 1.0));
            const int64_t x1 = std::min<int64_t>(R - 1, (int64_t)std::ceil(s.c[0] + s.r + 1.0));
            for (int64_t y = y0; y <= y1; ++y)
                for (int64_t x = x0; x <= x1; ++x)
                    if (std::fabs(sphereSdf(s, x + 0.5, y + 0.5, pz)) < 0.5)
                        setBit(bits, (uint64_t)x + (uint64_t)R * y + (uint64_t)R * R * z);
        }
        for (const Box& b : boxes) {
            if (pz < b.lo[2] - 1.0 || pz > b.lo[2] + b.e[2] + 1.0)
                continue;
            const int64_t y0 = std::max<int64_t>(0, (int64_t)std::floor(b.lo[1] - 1.0));
            const int64_t y1 = std::min<int64_t>(R - 1, (int64_t)std::ceil(b.lo[1] + b.e[1] + 1.0));
            const int64_t x0 = std::max<int64_t>(0, (int64_t)std::floor(b.lo[0] - 1.0));
            const int64_t x1 = std::min<int64_t>(R - 1, (int64_t)std::ceil(b.lo[0] + b.e[0] + 1.0));
            for (int64_t y = y0; y <= y1; ++y)
                for (int64_t x = x0; x <= x1; ++x)
                    if (std::fabs(boxSdf(b, x + 0.5, y + 0.5, pz)) < 0.5)
                        setBit(bits, (uint64_t)x + (uint64_t)R * y + (uint64_t)R * R * z);
        }
    });
}

// ---------------------------------------------------------------------------------------
// City (kind 2). Blocks of 4 x 4 lots of 64 x 64 voxels (256 x 256 per block, x and z
// horizontal, y up). Half of the blocks are originals: each lot holds a building with
// probability 0.85, footprint inset U{2..10} voxels from every lot edge, height U[32, 1024]
// (clipped to the grid), standing on the block's base (the ground height at the block centre).
// The other half copy a random earlier original under a random non-identity combination of
// mirror x, mirror z and swap x/z, on their own base, so exact translation, mirror and
// axis-permutation matches exist. Buildings are box shells (walls, floor, roof, one voxel
// thick) with window voids in the side walls: 4 voxels wide with pitch 8 along the wall (starting
// 2 voxels from the wall's anchored end, never within 2 voxels of either end) and 6 voxels high
// with pitch 8 (starting 4 voxels above the base, never within 2 voxels of the roof). A wall's
// anchored end is its low end in the original, so mirrored copies stay exact images. The ground
// is a 4-octave value-noise heightfield shell (heights 0-48 voxels).
// ---------------------------------------------------------------------------------------
struct Building {
    int32_t x0, x1, z0, z1; // world footprint [x0, x1) x [z0, z1)
    int32_t y0, y1; // [base, top)
    bool anchorLowX, anchorLowZ; // window pattern anchored at the low end of the wall along x / z
};

// Window void test of a wall position: `o` = offset from the wall's anchored end, `len` = wall
// length along that axis; `oy` = height above the base, `h` = building height.
inline bool windowAt(int32_t o, int32_t len, int32_t oy, int32_t h)
{
    const bool horiz = o >= 2 && o < len - 2 && ((o - 2) & 7) < 4;
    const bool vert = oy >= 4 && oy < h - 2 && ((oy - 4) & 7) < 6;
    return horiz && vert;
}

std::vector<Building> cityBuildings(uint32_t R, uint64_t seed, const std::vector<int32_t>& ground)
{
    constexpr int32_t kLot = 64, kBlock = 256;
    const int32_t nb = (int32_t)R / kBlock;
    struct LocalBuilding {
        int32_t x0, x1, z0, z1, h; // block-local footprint, height
    };
    std::vector<std::vector<LocalBuilding>> designs((size_t)nb * nb);
    std::vector<int32_t> originals;
    std::vector<Building> out;
    for (int32_t bz = 0; bz < nb; ++bz)
        for (int32_t bx = 0; bx < nb; ++bx) {
            const int32_t bid = bx + nb * bz;
            SplitMix64 rng { mix64(seed ^ 0xC17Cull ^ ((uint64_t)bid * 0x9E3779B97F4A7C15ull)) };
            const int32_t cx = bx * kBlock + kBlock / 2, cz = bz * kBlock + kBlock / 2;
            const int32_t base = ground[(size_t)cx + (size_t)R * cz];
            const bool copy = !originals.empty() && rng.uniform() < 0.5;
            uint32_t tf = 0; // bit 0: mirror x, bit 1: mirror z, bit 2: swap x/z
            int32_t src = bid;
            if (copy) {
                src = originals[(size_t)(rng.next() % originals.size())];
                tf = 1u + (uint32_t)(rng.next() % 7u);
            } else {
                for (int32_t lz = 0; lz < 4; ++lz)
                    for (int32_t lx = 0; lx < 4; ++lx) {
                        if (rng.uniform() >= 0.85)
                            continue;
                        LocalBuilding b;
                        b.x0 = lx * kLot + 2 + (int32_t)(rng.next() % 9u);
                        b.x1 = (lx + 1) * kLot - 2 - (int32_t)(rng.next() % 9u);
                        b.z0 = lz * kLot + 2 + (int32_t)(rng.next() % 9u);
                        b.z1 = (lz + 1) * kLot - 2 - (int32_t)(rng.next() % 9u);
                        b.h = 32 + (int32_t)(rng.uniform() * 993.0);
                        designs[(size_t)bid].push_back(b);
                    }
                originals.push_back(bid);
            }
            for (const LocalBuilding& lb : designs[(size_t)src]) {
                // block-local image of the source building under tf (mirrors, then the swap)
                int32_t x0 = lb.x0, x1 = lb.x1, z0 = lb.z0, z1 = lb.z1;
                bool lowX = true, lowZ = true;
                if (tf & 1u) {
                    const int32_t a = kBlock - x1, b = kBlock - x0;
                    x0 = a;
                    x1 = b;
                    lowX = false;
                }
                if (tf & 2u) {
                    const int32_t a = kBlock - z1, b = kBlock - z0;
                    z0 = a;
                    z1 = b;
                    lowZ = false;
                }
                if (tf & 4u) {
                    std::swap(x0, z0);
                    std::swap(x1, z1);
                    std::swap(lowX, lowZ);
                }
                Building w;
                w.x0 = bx * kBlock + x0;
                w.x1 = bx * kBlock + x1;
                w.z0 = bz * kBlock + z0;
                w.z1 = bz * kBlock + z1;
                w.y0 = base;
                w.y1 = std::min<int32_t>((int32_t)R, base + lb.h);
                w.anchorLowX = lowX;
                w.anchorLowZ = lowZ;
                if (w.y1 - w.y0 >= 3)
                    out.push_back(w);
            }
        }
    return out;
}

inline void setRange(uint8_t* bits, uint64_t rowBase, int32_t x0, int32_t x1)
{
    for (int32_t x = x0; x < x1; ++x)
        setBit(bits, rowBase + (uint64_t)x);
}

void rasterizeCity(uint8_t* bits, uint32_t R, const std::vector<Building>& buildings, int threads)
{
    parallelSlabs(R, threads, [&](uint32_t zu) {
        const int32_t z = (int32_t)zu;
        for (const Building& b : buildings) {
            if (z < b.z0 || z >= b.z1)
                continue;
            const int32_t h = b.y1 - b.y0;
            const bool frontBack = z == b.z0 || z == b.z1 - 1; // walls spanning x
            const int32_t lenX = b.x1 - b.x0, lenZ = b.z1 - b.z0;
            const int32_t oz = b.anchorLowZ ? z - b.z0 : b.z1 - 1 - z;
            for (int32_t y = b.y0; y < b.y1; ++y) {
                const uint64_t row = (uint64_t)R * (uint64_t)y + (uint64_t)R * R * (uint64_t)z;
                const int32_t oy = y - b.y0;
                if (y == b.y0 || y == b.y1 - 1) { // floor and roof
                    setRange(bits, row, b.x0, b.x1);
                    continue;
                }
                if (frontBack) {
                    for (int32_t x = b.x0; x < b.x1; ++x) {
                        const int32_t ox = b.anchorLowX ? x - b.x0 : b.x1 - 1 - x;
                        if (!windowAt(ox, lenX, oy, h))
                            setBit(bits, row + (uint64_t)x);
                    }
                    continue;
                }
                // side walls facing -x and +x: windows along z
                if (!windowAt(oz, lenZ, oy, h)) {
                    setBit(bits, row + (uint64_t)b.x0);
                    setBit(bits, row + (uint64_t)(b.x1 - 1));
                }
            }
        }
    });
}

} // namespace

// kind 0: 200 sphere shells (centres U[0,R)^3, radius U[4,40]) + 100 box shells (corner
//         U[0,R)^3, edges U[8,64]); voxel v is set when |SDF(v + 0.5)| < 0.5.
// kind 1: kind 0 + fractal value-noise heightfield shell (6 octaves, persistence 0.5,
//         heights 0-400 voxels).
// kind 2: the city described above (R >= 256).
bool generateSceneDense(uint32_t kind, uint32_t R, uint64_t seed, uint8_t* bits, int threads)
{
    if (R < 4 || (R & (R - 1)) != 0 || kind > 2)
        return false;
    std::memset(bits, 0, (size_t)((uint64_t)R * R * R / 8));
    if (kind == 2) {
        if (R < 256)
            return false;
        const std::vector<int32_t> ground = heightfieldTops(R, seed ^ 0x6E0D6E0D6E0D6E0Dull, 4, 48.0, threads);
        rasterizeTops(bits, R, ground, threads);
        rasterizeCity(bits, R, cityBuildings(R, seed, ground), threads);
        return true;
    }
    SplitMix64 rng { seed };
    std::vector<Sphere> spheres(200);
    for (Sphere& s : spheres) {
        for (int a = 0; a < 3; ++a)
            s.c[a] = rng.uniform() * R;
        s.r = 4.0 + 36.0 * rng.uniform();
    }
    std::vector<Box> boxes(100);
    for (Box& b : boxes) {
        for (int a = 0; a < 3; ++a)
            b.lo[a] = rng.uniform() * R;
        for (int a = 0; a < 3; ++a)
            b.e[a] = 8.0 + 56.0 * rng.uniform();
    }
    rasterizeShapes(bits, R, spheres, boxes, threads);
    if (kind == 1) {
        const std::vector<int32_t> top = heightfieldTops(R, seed ^ 0x5EED5EED5EED5EEDull, 6, 400.0, threads);
        rasterizeTops(bits, R, top, threads);
    }
    return true;
}

} // namespace tdag
