// Seeded synthetic image pairs for the benchmarks and tests.
//
// Scene: a procedural "world" texture - value noise (16-px lattice, +-20),
// axis-aligned rectangles and Gaussian blobs, counts scaled per area of image
// A (50 rectangles and 200 blobs per A) - covering A and the footprint of B.
// A is the world crop at the origin. B is the world seen through a random
// homography H (B -> A): rotation +-5 deg, scale 0.9-1.1, perspective terms
// +-1e-4, vertical shift +-5 % and a horizontal shift found by bisection so
// that the area of A covered by B hits the requested overlap fraction; B
// then gets N(0, 2^2) noise and a +-10 brightness offset. Variants: low
// texture (blobs / 10) and repetitive texture (a tiled 48x48 random patch).
// All randomness comes from one SplitMix64 stream seeded per pair, so every
// process produces the same bytes.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "cs_stitch.h"

namespace {

struct Rng {
    std::uint64_t s;
    std::uint64_t next() {
        std::uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    double normal() {
        const double u1 = std::max(uniform(), 1e-300), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct P2 {
    double x, y;
};

P2 apply(const double* h, double x, double y) {
    const double w = h[6] * x + h[7] * y + h[8];
    return {(h[0] * x + h[1] * y + h[2]) / w, (h[3] * x + h[4] * y + h[5]) / w};
}

// Area of polygon `poly` clipped to [0, W] x [0, H] (Sutherland-Hodgman).
double clipped_area(std::vector<P2> poly, double W, double H) {
    auto clip = [&](auto inside, auto cross) {
        std::vector<P2> out;
        for (std::size_t i = 0; i < poly.size(); ++i) {
            const P2 a = poly[i], b = poly[(i + 1) % poly.size()];
            const bool ia = inside(a), ib = inside(b);
            if (ia && ib) out.push_back(b);
            else if (ia && !ib) out.push_back(cross(a, b));
            else if (!ia && ib) {
                out.push_back(cross(a, b));
                out.push_back(b);
            }
        }
        poly = out;
    };
    auto xcut = [](double X) {
        return [X](P2 a, P2 b) { return P2{X, a.y + (b.y - a.y) * (X - a.x) / (b.x - a.x)}; };
    };
    auto ycut = [](double Y) {
        return [Y](P2 a, P2 b) { return P2{a.x + (b.x - a.x) * (Y - a.y) / (b.y - a.y), Y}; };
    };
    clip([](P2 p) { return p.x >= 0; }, xcut(0));
    if (poly.empty()) return 0;
    clip([W](P2 p) { return p.x <= W; }, xcut(W));
    if (poly.empty()) return 0;
    clip([](P2 p) { return p.y >= 0; }, ycut(0));
    if (poly.empty()) return 0;
    clip([H](P2 p) { return p.y <= H; }, ycut(H));
    double a = 0;
    for (std::size_t i = 0; i < poly.size(); ++i) {
        const P2 p = poly[i], q = poly[(i + 1) % poly.size()];
        a += p.x * q.y - q.x * p.y;
    }
    return std::fabs(a) * 0.5;
}

struct World {
    int x0, y0, w, h;
    std::vector<float> rgb;  // w*h*3
    float at(int x, int y, int c) const {
        x = std::clamp(x - x0, 0, w - 1);
        y = std::clamp(y - y0, 0, h - 1);
        return rgb[(static_cast<std::size_t>(y) * w + x) * 3 + c];
    }
};

void render_world(World& wd, Rng& rng, double a_area, int kind) {
    const double scale = static_cast<double>(wd.w) * wd.h / a_area;
    wd.rgb.assign(static_cast<std::size_t>(wd.w) * wd.h * 3, 128.0f);
    // Value noise on a 16-px lattice, amplitude 20, same for all channels.
    const int cell = 16;
    const int gw = wd.w / cell + 2, gh = wd.h / cell + 2;
    std::vector<float> lat(static_cast<std::size_t>(gw) * gh);
    for (auto& v : lat) v = static_cast<float>(rng.uniform(-20, 20));
    for (int y = 0; y < wd.h; ++y) {
        const int gy = y / cell;
        const float fy = static_cast<float>(y % cell) / cell;
        for (int x = 0; x < wd.w; ++x) {
            const int gx = x / cell;
            const float fx = static_cast<float>(x % cell) / cell;
            const float v = lat[gy * gw + gx] * (1 - fx) * (1 - fy) + lat[gy * gw + gx + 1] * fx * (1 - fy) +
                            lat[(gy + 1) * gw + gx] * (1 - fx) * fy + lat[(gy + 1) * gw + gx + 1] * fx * fy;
            float* p = &wd.rgb[(static_cast<std::size_t>(y) * wd.w + x) * 3];
            p[0] += v;
            p[1] += v;
            p[2] += v;
        }
    }
    if (kind == 2) {
        // Repetitive: tile a 48x48 random patch of rectangles and blobs over the world.
        const int T = 48;
        std::vector<float> tile(T * T * 3, 128.0f);
        for (int r = 0; r < 6; ++r) {
            const int x0 = static_cast<int>(rng.uniform(0, T)), y0 = static_cast<int>(rng.uniform(0, T));
            const int ww = 3 + static_cast<int>(rng.uniform(0, T / 3)), hh = 3 + static_cast<int>(rng.uniform(0, T / 3));
            const float col[3] = {static_cast<float>(rng.uniform(0, 255)), static_cast<float>(rng.uniform(0, 255)),
                                  static_cast<float>(rng.uniform(0, 255))};
            for (int y = y0; y < std::min(T, y0 + hh); ++y)
                for (int x = x0; x < std::min(T, x0 + ww); ++x)
                    for (int c = 0; c < 3; ++c) tile[(y * T + x) * 3 + c] = col[c];
        }
        for (int y = 0; y < wd.h; ++y)
            for (int x = 0; x < wd.w; ++x)
                for (int c = 0; c < 3; ++c)
                    wd.rgb[(static_cast<std::size_t>(y) * wd.w + x) * 3 + c] +=
                        tile[(((y + wd.y0) % T + T) % T * T + ((x + wd.x0) % T + T) % T) * 3 + c] - 128.0f;
        return;
    }
    const int rects = static_cast<int>(std::lround(50 * scale));
    for (int r = 0; r < rects; ++r) {
        const int x0 = static_cast<int>(rng.uniform(0, wd.w)), y0 = static_cast<int>(rng.uniform(0, wd.h));
        const int ww = 4 + static_cast<int>(rng.uniform(0, wd.w / 8.0)), hh = 4 + static_cast<int>(rng.uniform(0, wd.h / 8.0));
        const float col[3] = {static_cast<float>(rng.uniform(0, 255)), static_cast<float>(rng.uniform(0, 255)),
                              static_cast<float>(rng.uniform(0, 255))};
        for (int y = y0; y < std::min(wd.h, y0 + hh); ++y)
            for (int x = x0; x < std::min(wd.w, x0 + ww); ++x)
                for (int c = 0; c < 3; ++c) wd.rgb[(static_cast<std::size_t>(y) * wd.w + x) * 3 + c] = col[c];
    }
    const int blobs = static_cast<int>(std::lround((kind == 1 ? 20 : 200) * scale));
    for (int k = 0; k < blobs; ++k) {
        const double cx = rng.uniform(0, wd.w), cy = rng.uniform(0, wd.h), s = rng.uniform(4, 40);
        double amp[3];
        for (double& v : amp) v = (rng.uniform() < 0.5 ? -1 : 1) * rng.uniform(40, 120);
        const int r = static_cast<int>(3 * s) + 1;
        const int xa = std::max(0, static_cast<int>(cx) - r), xb = std::min(wd.w, static_cast<int>(cx) + r);
        const int ya = std::max(0, static_cast<int>(cy) - r), yb = std::min(wd.h, static_cast<int>(cy) + r);
        const double inv = -0.5 / (s * s);
        for (int y = ya; y < yb; ++y)
            for (int x = xa; x < xb; ++x) {
                const double g = std::exp(((x - cx) * (x - cx) + (y - cy) * (y - cy)) * inv);
                float* p = &wd.rgb[(static_cast<std::size_t>(y) * wd.w + x) * 3];
                for (int c = 0; c < 3; ++c) p[c] += static_cast<float>(amp[c] * g);
            }
    }
}

std::uint8_t clamp8(double v) { return static_cast<std::uint8_t>(std::clamp(std::lround(v), 0L, 255L)); }

}  // namespace

extern "C" int cs_synth_pair(std::uint64_t seed, int width, int height, double overlap_lo, double overlap_hi,
                             int kind, std::uint8_t* rgb_a, std::uint8_t* rgb_b, double* h_ba) {
    if (width < 8 || height < 8 || !rgb_a || !rgb_b || !(overlap_lo > 0) || overlap_hi > 1 || overlap_lo > overlap_hi ||
        kind < 0 || kind > 2)
        return CS_INVALID_ARGUMENT;
    Rng rng{seed};
    const double W = width, H = height;
    const double target = rng.uniform(overlap_lo, overlap_hi);
    const double th = rng.uniform(-5, 5) * 3.141592653589793 / 180.0, s = rng.uniform(0.9, 1.1);
    // Perspective terms ~1e-4; above 1920 px the term is scaled down so the
    // projective distortion across the frame stays that of a 1080p pair.
    const double pscale = std::min(1.0, 1920.0 / W);
    const double p1 = rng.uniform(-1e-4, 1e-4) * pscale, p2 = rng.uniform(-1e-4, 1e-4) * pscale;
    const double ty = rng.uniform(-0.05, 0.05) * H;
    double h[9] = {s * std::cos(th), -s * std::sin(th), 0, s * std::sin(th), s * std::cos(th), ty, p1, p2, 1};
    auto overlap = [&](double tx) {
        h[2] = tx;
        std::vector<P2> q = {apply(h, 0, 0), apply(h, W, 0), apply(h, W, H), apply(h, 0, H)};
        return clipped_area(q, W, H) / (W * H);
    };
    double lo = -0.5 * W, hi = 1.2 * W;  // overlap decreases as B moves right
    for (int it = 0; it < 80; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (overlap(mid) > target) lo = mid;
        else hi = mid;
    }
    overlap(0.5 * (lo + hi));
    // World covers A and the footprint of B in A coordinates.
    double minx = 0, miny = 0, maxx = W, maxy = H;
    for (P2 c : {apply(h, 0, 0), apply(h, W, 0), apply(h, W, H), apply(h, 0, H)}) {
        minx = std::min(minx, c.x);
        miny = std::min(miny, c.y);
        maxx = std::max(maxx, c.x);
        maxy = std::max(maxy, c.y);
    }
    World wd;
    wd.x0 = static_cast<int>(std::floor(minx)) - 2;
    wd.y0 = static_cast<int>(std::floor(miny)) - 2;
    wd.w = static_cast<int>(std::ceil(maxx)) + 3 - wd.x0;
    wd.h = static_cast<int>(std::ceil(maxy)) + 3 - wd.y0;
    render_world(wd, rng, W * H, kind);
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            for (int c = 0; c < 3; ++c)
                rgb_a[(static_cast<std::size_t>(y) * width + x) * 3 + c] = clamp8(wd.at(x, y, c));
    const double offset = rng.uniform(-10, 10);
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const P2 p = apply(h, x, y);
            const int x0 = static_cast<int>(std::floor(p.x)), y0 = static_cast<int>(std::floor(p.y));
            const double fx = p.x - x0, fy = p.y - y0;
            for (int c = 0; c < 3; ++c) {
                const double v = wd.at(x0, y0, c) * (1 - fx) * (1 - fy) + wd.at(x0 + 1, y0, c) * fx * (1 - fy) +
                                 wd.at(x0, y0 + 1, c) * (1 - fx) * fy + wd.at(x0 + 1, y0 + 1, c) * fx * fy;
                rgb_b[(static_cast<std::size_t>(y) * width + x) * 3 + c] = clamp8(v + 2.0 * rng.normal() + offset);
            }
        }
    if (h_ba) std::memcpy(h_ba, h, sizeof h);
    return CS_OK;
}

// A horizontal strip of n frames: frame k + 1 is seen from frame k through a
// random homography (rotation U[-1, 1] degrees, scale U[0.98, 1.02],
// perspective terms U[-1e-5, 1e-5] at 1080p, vertical shift U[-2, 2] % of
// the height; milder than a pair's because they compound along the chain)
// whose horizontal shift puts `overlap` of frame k under frame k + 1. Frame 0 is the world crop at the origin and the world covers
// every frame's footprint, with the pair generator's texture density per
// frame area; frames 1.. get the noise and brightness offset of a pair's B.
// h_adj[9k..9k+8] maps frame k + 1 to frame k.
extern "C" int cs_synth_strip(std::uint64_t seed, int n, int width, int height, double overlap, int kind,
                              std::uint8_t* frames, double* h_adj) {
    if (n < 1 || n > 64 || width < 8 || height < 8 || !frames || !(overlap > 0) || overlap >= 1 || kind < 0 ||
        kind > 2)
        return CS_INVALID_ARGUMENT;
    Rng rng{seed};
    const double W = width, H = height;
    const double pscale = std::min(1.0, 1920.0 / W);
    std::vector<std::array<double, 9>> to0(n);  // frame k -> frame 0
    to0[0] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int k = 1; k < n; ++k) {
        const double th = rng.uniform(-1, 1) * 3.141592653589793 / 180.0, s = rng.uniform(0.98, 1.02);
        const double p1 = rng.uniform(-1e-5, 1e-5) * pscale, p2 = rng.uniform(-1e-5, 1e-5) * pscale;
        const double ty = rng.uniform(-0.02, 0.02) * H;
        double h[9] = {s * std::cos(th), -s * std::sin(th), 0, s * std::sin(th), s * std::cos(th), ty, p1, p2, 1};
        auto covered = [&](double tx) {
            h[2] = tx;
            std::vector<P2> q = {apply(h, 0, 0), apply(h, W, 0), apply(h, W, H), apply(h, 0, H)};
            return clipped_area(q, W, H) / (W * H);
        };
        double lo = -0.5 * W, hi = 1.2 * W;
        for (int it = 0; it < 80; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (covered(mid) > overlap) lo = mid;
            else hi = mid;
        }
        covered(0.5 * (lo + hi));
        if (h_adj) std::memcpy(h_adj + 9 * (k - 1), h, sizeof h);
        const auto& g = to0[k - 1];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double acc = 0;
                for (int i = 0; i < 3; ++i) acc += g[3 * r + i] * h[3 * i + c];
                to0[k][3 * r + c] = acc;
            }
    }
    double minx = 0, miny = 0, maxx = W, maxy = H;
    for (int k = 1; k < n; ++k)
        for (P2 c : {apply(to0[k].data(), 0, 0), apply(to0[k].data(), W, 0), apply(to0[k].data(), W, H),
                     apply(to0[k].data(), 0, H)}) {
            minx = std::min(minx, c.x);
            miny = std::min(miny, c.y);
            maxx = std::max(maxx, c.x);
            maxy = std::max(maxy, c.y);
        }
    World wd;
    wd.x0 = static_cast<int>(std::floor(minx)) - 2;
    wd.y0 = static_cast<int>(std::floor(miny)) - 2;
    wd.w = static_cast<int>(std::ceil(maxx)) + 3 - wd.x0;
    wd.h = static_cast<int>(std::ceil(maxy)) + 3 - wd.y0;
    render_world(wd, rng, W * H, kind);
    const std::size_t frame_bytes = static_cast<std::size_t>(width) * height * 3;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            for (int c = 0; c < 3; ++c) frames[(static_cast<std::size_t>(y) * width + x) * 3 + c] = clamp8(wd.at(x, y, c));
    for (int k = 1; k < n; ++k) {
        std::uint8_t* out = frames + frame_bytes * k;
        const double offset = rng.uniform(-10, 10);
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                const P2 p = apply(to0[k].data(), x, y);
                const int x0 = static_cast<int>(std::floor(p.x)), y0 = static_cast<int>(std::floor(p.y));
                const double fx = p.x - x0, fy = p.y - y0;
                for (int c = 0; c < 3; ++c) {
                    const double v = wd.at(x0, y0, c) * (1 - fx) * (1 - fy) + wd.at(x0 + 1, y0, c) * fx * (1 - fy) +
                                     wd.at(x0, y0 + 1, c) * (1 - fx) * fy + wd.at(x0 + 1, y0 + 1, c) * fx * fy;
                    out[(static_cast<std::size_t>(y) * width + x) * 3 + c] = clamp8(v + 2.0 * rng.normal() + offset);
                }
            }
    }
    return CS_OK;
}
