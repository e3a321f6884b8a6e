// Host canvas geometry for S6, in extended (long double) precision.
//
// Contract (src/s6/geometry.hpp:10-44, semantics of src/s6/geometry.cpp):
// H maps B-original pixel centres to A-original ones. The canvas is the
// inclusive integer box around A's pixel centres and B's four projected
// corners, each bound snapped to the nearest integer when within tolerance;
// B's corners must all lie on one side of the horizon (denominators of one
// sign, each clear of zero); the canvas must fit the pixel budget; and the
// inverse map handed to the warp is computed in a well-conditioned frame
// (each image centred and scaled to unit size), normalised by its largest
// coefficient.
//
// Every value that reaches the output is produced by the same sequence of
// long double operations as the reference's (products summed left to right,
// divisions by the same divisors, the inverse by Gauss-Jordan elimination
// with scaled partial pivoting), so canvases and inverse maps are bit-equal
// to the reference's; only the organisation differs.
#include <algorithm>
#include <cmath>
#include <limits>
#include <utility>

#include "stitch/stages.hpp"

namespace stitch::s6 {
namespace {

using Real = long double;

// Row-major 3x3 matrix with value semantics.
struct Mat3 {
    Real e[9]{};

    Real& at(int r, int c) { return e[3 * r + c]; }
    Real at(int r, int c) const { return e[3 * r + c]; }

    // [[s, 0, tx], [0, s, ty], [0, 0, 1]]: an isotropic scale and shift.
    static Mat3 similarity(Real s, Real tx, Real ty) {
        Mat3 m;
        m.at(0, 0) = m.at(1, 1) = s;
        m.at(0, 2) = tx;
        m.at(1, 2) = ty;
        m.at(2, 2) = 1;
        return m;
    }

    friend Mat3 operator*(const Mat3& a, const Mat3& b) {
        Mat3 p;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                Real acc = 0;
                for (int k = 0; k < 3; ++k) acc += a.at(r, k) * b.at(k, c);
                p.at(r, c) = acc;
            }
        return p;
    }

    // Divides by the largest magnitude; false for a zero or non-finite matrix.
    bool normalize() {
        Real peak = 0;
        for (Real v : e) {
            if (!std::isfinite(v)) return false;
            peak = std::max(peak, std::abs(v));
        }
        if (peak == 0) return false;
        for (Real& v : e) v /= peak;
        return true;
    }

    // Maximum absolute row sum (the infinity norm).
    Real row_sum_norm() const {
        Real best = 0;
        for (int r = 0; r < 3; ++r) best = std::max(best, std::abs(at(r, 0)) + std::abs(at(r, 1)) + std::abs(at(r, 2)));
        return best;
    }
};

// Inverse by Gauss-Jordan elimination on the working copy w while the same
// row operations turn x (initially I) into the inverse. The pivot of each
// column is the row with the largest |entry| relative to that row's
// original largest magnitude. Columns of w already reduced to unit vectors
// are never read again, so only the columns still to be reduced are updated.
// Fails when a pivot is not clearly non-zero or when the condition estimate
// ||m|| ||m^-1|| reaches 1 / tolerance.
bool invert_checked(const Mat3& m, Real tolerance, Mat3* inverse) {
    Mat3 w = m, x = Mat3::similarity(1, 0, 0);
    Real row_scale[3];
    for (int r = 0; r < 3; ++r) {
        row_scale[r] = std::max({std::abs(w.at(r, 0)), std::abs(w.at(r, 1)), std::abs(w.at(r, 2))});
        if (row_scale[r] == 0) return false;
    }
    for (int k = 0; k < 3; ++k) {
        int pivot = k;
        Real score = 0;
        for (int r = k; r < 3; ++r) {
            const Real s = std::abs(w.at(r, k)) / row_scale[r];
            if (s > score) {
                score = s;
                pivot = r;
            }
        }
        if (!(std::isfinite(score) && score > tolerance)) return false;
        if (pivot != k) {
            for (int c = 0; c < 3; ++c) {
                std::swap(w.at(pivot, c), w.at(k, c));
                std::swap(x.at(pivot, c), x.at(k, c));
            }
            std::swap(row_scale[pivot], row_scale[k]);
        }
        const Real d = w.at(k, k);
        for (int c = k; c < 3; ++c) w.at(k, c) /= d;
        for (int c = 0; c < 3; ++c) x.at(k, c) /= d;
        for (int r = 0; r < 3; ++r) {
            if (r == k) continue;
            const Real f = w.at(r, k);
            for (int c = k; c < 3; ++c) w.at(r, c) -= f * w.at(k, c);
            for (int c = 0; c < 3; ++c) x.at(r, c) -= f * x.at(k, c);
        }
    }
    for (Real v : x.e)
        if (!std::isfinite(v)) return false;
    const Real condition = m.row_sum_norm() * x.row_sum_norm();
    if (!(std::isfinite(condition) && condition * tolerance < 1)) return false;
    *inverse = x;
    return true;
}

// Inclusive box of pixel centres, grown point by point.
struct Box {
    Real x0, y0, x1, y1;
    void include(Real x, Real y) {
        x0 = std::min(x0, x);
        y0 = std::min(y0, y);
        x1 = std::max(x1, x);
        y1 = std::max(y1, y);
    }
};

// x moved onto the nearest integer when it lies within the snap tolerance of
// it (at least 16 double ulps of |x|, so coordinates that are integers up to
// rounding do not widen the canvas by a pixel).
Real snapped(Real x, double tolerance) {
    const Real nearest = std::round(x);
    const Real ulps = 16 * std::numeric_limits<double>::epsilon() * std::max(Real(1), std::abs(x));
    return std::abs(x - nearest) <= std::max(Real(tolerance), ulps) ? nearest : x;
}

bool usable(const GeometryPolicy& p) {
    const auto open_unit = [](double v) { return std::isfinite(v) && v > 0 && v < 1; };
    return open_unit(p.denominator_tolerance) && open_unit(p.inverse_tolerance) && std::isfinite(p.snap_tolerance) &&
           p.snap_tolerance >= 0 && p.snap_tolerance < 0.5 && p.max_canvas_pixels > 0;
}

}  // namespace

Status prepare_geometry(int wa, int ha, int wb, int hb, const double* h, const GeometryPolicy& policy, Geometry* out) {
    if (!out) return Status::INVALID_ARGUMENT;
    *out = Geometry{};
    if (!h || wa <= 0 || ha <= 0 || wb <= 0 || hb <= 0 || !usable(policy)) return Status::INVALID_ARGUMENT;
    Mat3 H;
    std::copy(h, h + 9, H.e);
    if (!H.normalize()) return Status::DEGENERATE;

    // B's corners through H. The horizon test compares each denominator with
    // the largest value the denominator row can reach over B.
    const Real bx = Real(wb) - 1, by = Real(hb) - 1;
    const Real reach = std::abs(H.at(2, 0)) * bx + std::abs(H.at(2, 1)) * by + std::abs(H.at(2, 2));
    if (!(reach > 0) || !std::isfinite(reach)) return Status::DEGENERATE;
    const Real clearance = policy.denominator_tolerance * reach;
    Box box{0, 0, Real(wa) - 1, Real(ha) - 1};
    bool ahead = false, behind = false;
    for (const auto& [u, v] : {std::pair<Real, Real>{0, 0}, {bx, 0}, {0, by}, {bx, by}}) {
        const Real z = H.at(2, 0) * u + H.at(2, 1) * v + H.at(2, 2);
        if (!std::isfinite(z) || std::abs(z) <= clearance) return Status::DEGENERATE;
        (z > 0 ? ahead : behind) = true;
        if (ahead && behind) return Status::DEGENERATE;
        const Real x = (H.at(0, 0) * u + H.at(0, 1) * v + H.at(0, 2)) / z;
        const Real y = (H.at(1, 0) * u + H.at(1, 1) * v + H.at(1, 2)) / z;
        if (!std::isfinite(x) || !std::isfinite(y)) return Status::DEGENERATE;
        box.include(x, y);
    }
    const Real left = std::floor(snapped(box.x0, policy.snap_tolerance));
    const Real top = std::floor(snapped(box.y0, policy.snap_tolerance));
    const Real right = std::ceil(snapped(box.x1, policy.snap_tolerance));
    const Real bottom = std::ceil(snapped(box.y1, policy.snap_tolerance));
    const Real width = right - left + 1, height = bottom - top + 1;

    // The canvas must be addressable with int coordinates and fit the budget.
    constexpr Real int_lo = std::numeric_limits<int>::min(), int_hi = std::numeric_limits<int>::max();
    const bool in_int = left >= int_lo && left <= int_hi && top >= int_lo && top <= int_hi;
    if (!in_int || !std::isfinite(width) || !std::isfinite(height) || width < 1 || height < 1 || width > int_hi ||
        height > int_hi)
        return Status::INSUFFICIENT_WORKSPACE;
    const auto cols = static_cast<std::size_t>(width), rows = static_cast<std::size_t>(height);
    if (cols > std::numeric_limits<std::size_t>::max() / rows || cols > policy.max_canvas_pixels / rows)
        return Status::INSUFFICIENT_WORKSPACE;

    // Conditioning frame: A-original -> unit A (centred, scaled by its larger
    // extent) and unit B -> B-original. H in that frame must be safely
    // invertible; the inverse goes back to original coordinates.
    const Real extent_a = std::max({Real(1), Real(wa) - 1, Real(ha) - 1});
    const Real extent_b = std::max({Real(1), bx, by});
    const Real cxa = (Real(wa) - 1) / 2, cya = (Real(ha) - 1) / 2;
    const Mat3 to_unit_a = Mat3::similarity(1 / extent_a, -cxa / extent_a, -cya / extent_a);
    const Mat3 from_unit_b = Mat3::similarity(extent_b, bx / 2, by / 2);
    Mat3 framed = to_unit_a * H * from_unit_b;
    Mat3 framed_inverse;
    if (!framed.normalize() || !invert_checked(framed, policy.inverse_tolerance, &framed_inverse))
        return Status::DEGENERATE;
    Mat3 inverse = from_unit_b * framed_inverse * to_unit_a;
    if (!inverse.normalize()) return Status::DEGENERATE;

    Geometry g{};
    g.width = static_cast<int>(width);
    g.height = static_cast<int>(height);
    g.origin_x = static_cast<int>(left);
    g.origin_y = static_cast<int>(top);
    g.width_a = wa;
    g.height_a = ha;
    g.width_b = wb;
    g.height_b = hb;
    g.denominator_tolerance = policy.denominator_tolerance;
    g.snap_tolerance = policy.snap_tolerance;
    for (int i = 0; i < 9; ++i) {
        g.inverse[i] = static_cast<double>(inverse.e[i]);
        if (!std::isfinite(g.inverse[i])) return Status::DEGENERATE;
    }
    *out = g;
    return Status::OK;
}

// Chain canvas (no counterpart in the reference, which stitches pairs).
// Frame k reaches frame 0 through G_k = G_{k-1} h_adj[k-1], kept normalised
// by its largest coefficient (FP64, products summed left to right). Each
// frame's geometry against frame 0 comes from prepare_geometry with an
// unbounded budget; the chain canvas is the union of those canvases (each
// already holds frame 0), and only the union must fit the policy's budget.
Status prepare_chain_geometry(int n, const int* widths, const int* heights, const double* h_adj,
                              const GeometryPolicy& policy, Geometry* out) {
    if (!out || n < 1 || n > kMaxChainFrames || !widths || !heights || (n > 1 && !h_adj))
        return Status::INVALID_ARGUMENT;
    for (int k = 0; k < n; ++k) {
        out[k] = Geometry{};
        if (widths[k] <= 0 || heights[k] <= 0) return Status::INVALID_ARGUMENT;
    }
    GeometryPolicy unbounded = policy;
    unbounded.max_canvas_pixels = std::numeric_limits<std::size_t>::max();
    double to0[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    long long left = 0, top = 0, right = widths[0], bottom = heights[0];  // half-open box in frame-0 pixels
    for (int k = 1; k < n; ++k) {
        const double* step = h_adj + 9 * (k - 1);
        double next[9];
        double peak = 0;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double acc = 0;
                for (int i = 0; i < 3; ++i) acc += to0[3 * r + i] * step[3 * i + c];
                next[3 * r + c] = acc;
                if (!std::isfinite(acc)) return Status::DEGENERATE;
                peak = std::max(peak, std::fabs(acc));
            }
        if (peak == 0) return Status::DEGENERATE;
        for (int i = 0; i < 9; ++i) to0[i] = next[i] / peak;
        const Status s = prepare_geometry(widths[0], heights[0], widths[k], heights[k], to0, unbounded, &out[k]);
        if (s != Status::OK) return s;
        left = std::min<long long>(left, out[k].origin_x);
        top = std::min<long long>(top, out[k].origin_y);
        right = std::max<long long>(right, static_cast<long long>(out[k].origin_x) + out[k].width);
        bottom = std::max<long long>(bottom, static_cast<long long>(out[k].origin_y) + out[k].height);
    }
    const long long cols = right - left, rows = bottom - top;
    if (cols > std::numeric_limits<int>::max() || rows > std::numeric_limits<int>::max() ||
        static_cast<std::size_t>(cols) > policy.max_canvas_pixels / static_cast<std::size_t>(rows)) {
        for (int k = 0; k < n; ++k) out[k] = Geometry{};
        return Status::INSUFFICIENT_WORKSPACE;
    }
    for (int k = 0; k < n; ++k) {
        Geometry& g = out[k];
        g.width = static_cast<int>(cols);
        g.height = static_cast<int>(rows);
        g.origin_x = static_cast<int>(left);
        g.origin_y = static_cast<int>(top);
        g.width_a = widths[0];
        g.height_a = heights[0];
        g.width_b = widths[k];
        g.height_b = heights[k];
        g.denominator_tolerance = policy.denominator_tolerance;
        g.snap_tolerance = policy.snap_tolerance;
        if (k == 0) {
            std::fill(std::begin(g.inverse), std::end(g.inverse), 0.0);
            g.inverse[0] = g.inverse[4] = g.inverse[8] = 1;
        }
    }
    return Status::OK;
}

}  // namespace stitch::s6
