// Host canvas geometry for S6 in extended (long double) precision.
// Semantics: src/s6/geometry.cpp:106-200 - normalise H, require one
// denominator sign at B's corners, snapped integer bounding box of A and
// H(B), canvas budget, conditioning-checked inverse in normalised coordinates.
#include <algorithm>
#include <cmath>
#include <limits>
#include <utility>

#include "stitch/stages.hpp"

namespace stitch::s6 {
namespace {

using Real = long double;

bool valid_policy(const GeometryPolicy& p) {
    return std::isfinite(p.denominator_tolerance) && p.denominator_tolerance > 0 && p.denominator_tolerance < 1 &&
           std::isfinite(p.inverse_tolerance) && p.inverse_tolerance > 0 && p.inverse_tolerance < 1 &&
           std::isfinite(p.snap_tolerance) && p.snap_tolerance >= 0 && p.snap_tolerance < 0.5 &&
           p.max_canvas_pixels > 0;
}

void mul(const Real* a, const Real* b, Real* out) {
    Real t[9] = {};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) t[3 * r + c] += a[3 * r + k] * b[3 * k + c];
    std::copy(t, t + 9, out);
}

bool scale_to_unit(Real* m) {
    Real largest = 0;
    for (int i = 0; i < 9; ++i) {
        if (!std::isfinite(m[i])) return false;
        largest = std::max(largest, std::abs(m[i]));
    }
    if (largest == 0) return false;
    for (int i = 0; i < 9; ++i) m[i] /= largest;
    return true;
}

Real row_norm(const Real* m) {
    Real n = 0;
    for (int r = 0; r < 3; ++r) n = std::max(n, std::abs(m[3 * r]) + std::abs(m[3 * r + 1]) + std::abs(m[3 * r + 2]));
    return n;
}

// Gauss-Jordan with scaled partial pivoting; rejects ill-conditioned input.
bool invert(const Real* in, Real tol, Real* inv) {
    Real a[3][6] = {};
    Real scale[3] = {};
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            a[r][c] = in[3 * r + c];
            scale[r] = std::max(scale[r], std::abs(a[r][c]));
        }
        if (scale[r] == 0) return false;
        a[r][r + 3] = 1;
    }
    for (int col = 0; col < 3; ++col) {
        int pivot = col;
        Real best = 0;
        for (int r = col; r < 3; ++r) {
            const Real score = std::abs(a[r][col]) / scale[r];
            if (score > best) {
                best = score;
                pivot = r;
            }
        }
        if (!std::isfinite(best) || best <= tol) return false;
        if (pivot != col) {
            for (int c = 0; c < 6; ++c) std::swap(a[pivot][c], a[col][c]);
            std::swap(scale[pivot], scale[col]);
        }
        const Real d = a[col][col];
        for (int c = 0; c < 6; ++c) a[col][c] /= d;
        for (int r = 0; r < 3; ++r) {
            if (r == col) continue;
            const Real f = a[r][col];
            for (int c = 0; c < 6; ++c) a[r][c] -= f * a[col][c];
        }
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            if (!std::isfinite(a[r][c + 3])) return false;
            inv[3 * r + c] = a[r][c + 3];
        }
    const Real cond = row_norm(in) * row_norm(inv);
    return std::isfinite(cond) && cond * tol < 1;
}

Real snap(Real x, const GeometryPolicy& p) {
    const Real rounded = std::round(x);
    const Real floor16 = 16 * std::numeric_limits<double>::epsilon() * std::max(Real(1), std::abs(x));
    const Real tol = std::max(Real(p.snap_tolerance), floor16);
    return std::abs(x - rounded) <= tol ? rounded : x;
}

}  // namespace

Status prepare_geometry(int wa, int ha, int wb, int hb, const double* h, const GeometryPolicy& policy, Geometry* out) {
    if (!out) return Status::INVALID_ARGUMENT;
    Real m[9] = {};
    if (h) std::copy(h, h + 9, m);
    *out = Geometry{};
    if (!h || wa <= 0 || ha <= 0 || wb <= 0 || hb <= 0 || !valid_policy(policy)) return Status::INVALID_ARGUMENT;
    if (!scale_to_unit(m)) return Status::DEGENERATE;
    const Real bx = Real(wb) - 1, by = Real(hb) - 1;
    const Real dscale = std::abs(m[6]) * bx + std::abs(m[7]) * by + std::abs(m[8]);
    if (!(dscale > 0) || !std::isfinite(dscale)) return Status::DEGENERATE;
    const Real dfloor = policy.denominator_tolerance * dscale;
    Real xmin = 0, ymin = 0, xmax = Real(wa) - 1, ymax = Real(ha) - 1;
    int sign0 = 0;
    for (int corner = 0; corner < 4; ++corner) {
        const Real x = (corner & 1) ? bx : 0, y = (corner & 2) ? by : 0;
        const Real w = m[6] * x + m[7] * y + m[8];
        if (!std::isfinite(w) || std::abs(w) <= dfloor) return Status::DEGENERATE;
        const int sign = w > 0 ? 1 : -1;
        if (sign0 && sign != sign0) return Status::DEGENERATE;
        sign0 = sign;
        const Real px = (m[0] * x + m[1] * y + m[2]) / w;
        const Real py = (m[3] * x + m[4] * y + m[5]) / w;
        if (!std::isfinite(px) || !std::isfinite(py)) return Status::DEGENERATE;
        xmin = std::min(xmin, px);
        ymin = std::min(ymin, py);
        xmax = std::max(xmax, px);
        ymax = std::max(ymax, py);
    }
    xmin = std::floor(snap(xmin, policy));
    ymin = std::floor(snap(ymin, policy));
    xmax = std::ceil(snap(xmax, policy));
    ymax = std::ceil(snap(ymax, policy));
    const Real width = xmax - xmin + 1, height = ymax - ymin + 1;
    constexpr int imax = std::numeric_limits<int>::max(), imin = std::numeric_limits<int>::min();
    if (xmin < imin || ymin < imin || xmin > imax || ymin > imax || !std::isfinite(width) || !std::isfinite(height) ||
        width < 1 || height < 1 || width > imax || height > imax)
        return Status::INSUFFICIENT_WORKSPACE;
    const auto cw = static_cast<std::size_t>(width), ch = static_cast<std::size_t>(height);
    if (cw > std::numeric_limits<std::size_t>::max() / ch || cw > policy.max_canvas_pixels / ch)
        return Status::INSUFFICIENT_WORKSPACE;
    const Real sa = std::max({Real(1), Real(wa) - 1, Real(ha) - 1});
    const Real sb = std::max({Real(1), bx, by});
    const Real axc = (Real(wa) - 1) / 2, ayc = (Real(ha) - 1) / 2;
    const Real na[9] = {1 / sa, 0, -axc / sa, 0, 1 / sa, -ayc / sa, 0, 0, 1};
    const Real nbi[9] = {sb, 0, bx / 2, 0, sb, by / 2, 0, 0, 1};
    Real t[9], nh[9], ninv[9], oinv[9];
    mul(na, m, t);
    mul(t, nbi, nh);
    if (!scale_to_unit(nh) || !invert(nh, policy.inverse_tolerance, ninv)) return Status::DEGENERATE;
    mul(nbi, ninv, t);
    mul(t, na, oinv);
    if (!scale_to_unit(oinv)) return Status::DEGENERATE;
    Geometry g{};
    g.width = static_cast<int>(width);
    g.height = static_cast<int>(height);
    g.origin_x = static_cast<int>(xmin);
    g.origin_y = static_cast<int>(ymin);
    g.width_a = wa;
    g.height_a = ha;
    g.width_b = wb;
    g.height_b = hb;
    g.denominator_tolerance = policy.denominator_tolerance;
    g.snap_tolerance = policy.snap_tolerance;
    for (int i = 0; i < 9; ++i) {
        g.inverse[i] = static_cast<double>(oinv[i]);
        if (!std::isfinite(g.inverse[i])) return Status::DEGENERATE;
    }
    *out = g;
    return Status::OK;
}

}  // namespace stitch::s6
