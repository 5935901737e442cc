// Synthetic input: a closed lathe-revolved vessel (the reference's dataset is
// archaeological vessels; its meshes are not shipped).  A Catmull-Rom radius
// profile r(h) from seeded control points is revolved about the y axis into an
// outer wall, an inner wall offset by the wall thickness (3% of the height),
// an outer bottom disk, an inner bottom disk and a rim annulus, which together
// bound a closed, consistently oriented surface.  The mesh is then normalised
// like CADModel::load (MF/Source/Graphics/Core/CADModel.cpp:148-156): centred on
// its box and scaled so the largest half extent is 1.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace vf {

static double unit(uint64_t seed, uint32_t k) { return (double)hash32(seed, k) / 4294967296.0; }

static double catmull_rom(const std::vector<double>& c, double t)
{
    // uniform Catmull-Rom through the control values, end points repeated
    const int n = (int)c.size();
    double s = t * (n - 1);
    int i = (int)std::floor(s);
    if (i >= n - 1) i = n - 2;
    if (i < 0) i = 0;
    double u = s - i;
    double p0 = c[i > 0 ? i - 1 : 0], p1 = c[i], p2 = c[i + 1], p3 = c[i + 2 < n ? i + 2 : n - 1];
    return 0.5 * ((2 * p1) + (-p0 + p2) * u + (2 * p0 - 5 * p1 + 4 * p2 - p3) * u * u + (-p0 + 3 * p1 - 3 * p2 + p3) * u * u * u);
}

}  // namespace vf

extern "C" int vf_vessel_mesh(uint64_t seed, uint32_t R, uint32_t H, float* xyz, uint32_t* nv, uint32_t* tri, uint32_t* nt)
{
    return vf::guard([&] {
        if (R < 3 || H < 1 || R > 100000 || H > 100000) vf::fail(VF_ERR_INVALID_ARG, "radial >= 3 and height segments >= 1 required");
        const uint64_t nverts = 2ull * (H + 1) * R + 2;
        const uint64_t ntris = 4ull * R * H + 4ull * R;
        if (ntris > 0xFFFFFFFFull / 3) vf::fail(VF_ERR_INVALID_ARG, "mesh too large");
        if (nv) *nv = (uint32_t)nverts;
        if (nt) *nt = (uint32_t)ntris;
        if (!xyz && !tri) return;
        if (!xyz || !tri) vf::fail(VF_ERR_INVALID_ARG, "pass both buffers or neither");

        const int nctrl = 6 + (int)(vf::hash32(seed, 0) % 5u);
        std::vector<double> ctrl(nctrl);
        for (int k = 0; k < nctrl; ++k) ctrl[k] = 0.18 + 0.32 * vf::unit(seed, 1 + k);
        const double t = 0.03;  // wall thickness, 3% of the unit height
        auto radius = [&](double h) {
            double r = vf::catmull_rom(ctrl, h);
            return r < t + 0.06 ? t + 0.06 : r;
        };
        std::vector<double> P(3 * nverts);
        const double two_pi = 6.283185307179586;
        auto put = [&](uint64_t v, double r, double h, uint32_t j) {
            double a = two_pi * j / R;
            P[3 * v] = r * std::cos(a);
            P[3 * v + 1] = h;
            P[3 * v + 2] = r * std::sin(a);
        };
        // outer rings 0..H, then inner rings 0..H, then the two bottom centres
        for (uint32_t i = 0; i <= H; ++i) {
            double h = (double)i / H;
            for (uint32_t j = 0; j < R; ++j) put((uint64_t)i * R + j, radius(h), h, j);
        }
        const uint64_t inner0 = (uint64_t)(H + 1) * R;
        for (uint32_t i = 0; i <= H; ++i) {
            double h = t + (1.0 - t) * i / H;
            for (uint32_t j = 0; j < R; ++j) put(inner0 + (uint64_t)i * R + j, radius(h) - t, h, j);
        }
        const uint64_t cb = 2ull * (H + 1) * R, ci = cb + 1;
        P[3 * cb] = 0; P[3 * cb + 1] = 0; P[3 * cb + 2] = 0;
        P[3 * ci] = 0; P[3 * ci + 1] = t; P[3 * ci + 2] = 0;

        uint64_t f = 0;
        auto tri3 = [&](uint64_t a, uint64_t b, uint64_t c) {
            tri[3 * f] = (uint32_t)a;
            tri[3 * f + 1] = (uint32_t)b;
            tri[3 * f + 2] = (uint32_t)c;
            ++f;
        };
        auto O = [&](uint32_t i, uint32_t j) { return (uint64_t)i * R + (j % R); };
        auto I = [&](uint32_t i, uint32_t j) { return inner0 + (uint64_t)i * R + (j % R); };
        for (uint32_t i = 0; i < H; ++i)
            for (uint32_t j = 0; j < R; ++j) {
                tri3(O(i, j), O(i + 1, j), O(i + 1, j + 1));  // outer wall, normal outwards
                tri3(O(i, j), O(i + 1, j + 1), O(i, j + 1));
                tri3(I(i, j), I(i + 1, j + 1), I(i + 1, j));  // inner wall, normal into the cavity
                tri3(I(i, j), I(i, j + 1), I(i + 1, j + 1));
            }
        for (uint32_t j = 0; j < R; ++j) {
            tri3(cb, O(0, j), O(0, j + 1));      // outer bottom, normal down
            tri3(ci, I(0, j + 1), I(0, j));      // inner bottom, normal up into the cavity
            tri3(O(H, j), I(H, j), I(H, j + 1)); // rim annulus, normal up
            tri3(O(H, j), I(H, j + 1), O(H, j + 1));
        }
        // normalise: centre on the box, largest half extent 1
        double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
        for (uint64_t v = 0; v < nverts; ++v)
            for (int q = 0; q < 3; ++q) {
                lo[q] = std::fmin(lo[q], P[3 * v + q]);
                hi[q] = std::fmax(hi[q], P[3 * v + q]);
            }
        double half = 0;
        for (int q = 0; q < 3; ++q) half = std::fmax(half, 0.5 * (hi[q] - lo[q]));
        for (uint64_t v = 0; v < nverts; ++v)
            for (int q = 0; q < 3; ++q) xyz[3 * v + q] = (float)((P[3 * v + q] - 0.5 * (lo[q] + hi[q])) / half);
    });
}
