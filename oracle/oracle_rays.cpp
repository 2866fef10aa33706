// TEST INFRASTRUCTURE ONLY: first-hit ray casting through a dense voxel grid, the definition the
// device traversal of the .tdag16 (the renderer's use of the format,
// renderer/src/dags/transform_dag/transform_dag.cu:5-41) is checked against. A ray's hit is the
// first occupied voxel whose box the ray crosses (with positive length) inside [tmin, tmax];
// its distance is the ray's entry into that box. Double precision, Amanatides-Woo voxel walk.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

#include "oracle.h"

namespace {

bool occupied(const uint8_t* bits, uint32_t R, int x, int y, int z)
{
    const uint64_t i = (uint64_t)x + (uint64_t)R * ((uint64_t)y + (uint64_t)R * (uint64_t)z);
    return (bits[i >> 3] >> (i & 7)) & 1u;
}

} // namespace

extern "C" int orc_ray_march(const uint8_t* bits, uint32_t R, const float* rays, uint64_t n, int32_t* outHit, double* outT)
{
    const double inf = std::numeric_limits<double>::infinity();
    for (uint64_t r = 0; r < n; ++r) {
        const float* in = rays + r * 8;
        const double o[3] = { in[0], in[1], in[2] }, d[3] = { in[4], in[5], in[6] };
        int32_t* hit = outHit + r * 3;
        hit[0] = hit[1] = hit[2] = -1;
        outT[r] = 0.0;
        // Interval of the ray inside the volume.
        double t0 = in[3], t1 = in[7];
        bool none = false;
        for (int a = 0; a < 3; ++a) {
            if (d[a] == 0.0) {
                if (o[a] < 0.0 || o[a] >= (double)R)
                    none = true;
                continue;
            }
            double x0 = (0.0 - o[a]) / d[a], x1 = ((double)R - o[a]) / d[a];
            if (x0 > x1)
                std::swap(x0, x1);
            t0 = std::max(t0, x0);
            t1 = std::min(t1, x1);
        }
        if (none || !(t0 < t1))
            continue;
        int v[3], step[3];
        double tNext[3], tDelta[3];
        for (int a = 0; a < 3; ++a) {
            const double p = o[a] + t0 * d[a];
            v[a] = std::min(std::max((int)std::floor(p), 0), (int)R - 1);
            if (d[a] > 0.0) {
                step[a] = 1;
                tNext[a] = ((double)(v[a] + 1) - o[a]) / d[a];
                tDelta[a] = 1.0 / d[a];
            } else if (d[a] < 0.0) {
                step[a] = -1;
                tNext[a] = ((double)v[a] - o[a]) / d[a];
                tDelta[a] = -1.0 / d[a];
            } else {
                step[a] = 0;
                tNext[a] = inf;
                tDelta[a] = inf;
            }
        }
        double t = t0;
        for (;;) {
            const double tExit = std::min(tNext[0], std::min(tNext[1], tNext[2]));
            // The ray crosses voxel v over [t, min(tExit, t1)].
            if (std::min(tExit, t1) > t && occupied(bits, R, v[0], v[1], v[2])) {
                hit[0] = v[0];
                hit[1] = v[1];
                hit[2] = v[2];
                outT[r] = t;
                break;
            }
            if (tExit >= t1)
                break;
            const int a = tNext[0] == tExit ? 0 : (tNext[1] == tExit ? 1 : 2);
            v[a] += step[a];
            if (v[a] < 0 || v[a] >= (int)R)
                break;
            t = tExit;
            tNext[a] += tDelta[a];
        }
    }
    return 0;
}
