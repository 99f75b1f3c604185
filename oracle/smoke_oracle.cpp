// CPU oracle of the smoke step -- TEST INFRASTRUCTURE ONLY (see smoke_oracle.hpp).
//
// Restates the reference step (backend-cuda.cu, "B") stage by stage.  Each
// function cites the reference lines it follows.  Two places deliberately fix
// semantics the reference leaves racy:
//   * enforce / project: a domain OUTFLOW face copies the adjacent face AFTER
//     every interior face of the pass has been updated (B:1023-1024 vs 1063,
//     B:1435-1446 vs 1480 race in the reference);
//   * when an axis has a single cell, both domain faces of a line read their
//     neighbour before either is written.
#include "smoke_oracle.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using Flow = SmokeSimulationFlowBoundaryConfig;
using Scal = SmokeSimulationScalarBoundaryConfig;
using FlowFace = SmokeSimulationFlowBoundaryFaceDesc;
using ScalFace = SmokeSimulationScalarBoundaryFaceDesc;

constexpr uint32_t kFreeSlip = SMOKE_SIMULATION_FLOW_BOUNDARY_FREE_SLIP_WALL;
constexpr uint32_t kOutflow = SMOKE_SIMULATION_FLOW_BOUNDARY_OUTFLOW;
constexpr uint32_t kPeriodic = SMOKE_SIMULATION_FLOW_BOUNDARY_PERIODIC;
constexpr uint32_t kZeroFlux = SMOKE_SIMULATION_SCALAR_BOUNDARY_ZERO_FLUX;
constexpr uint32_t kScalarPeriodic = SMOKE_SIMULATION_SCALAR_BOUNDARY_PERIODIC;

struct Geo {
    int n[3];
    double h;
    const Flow* flow;
    const Scal* density_bc;
    const Scal* temperature_bc;
    bool flow_periodic[3];
    int64_t cells;
    int64_t faces[3];
};

const FlowFace& flow_face(const Flow& f, int axis, int side) {
    const FlowFace* faces[6] = {&f.x_minus, &f.x_plus, &f.y_minus, &f.y_plus, &f.z_minus, &f.z_plus};
    return *faces[axis * 2 + side];
}

const ScalFace& scal_face(const Scal& f, int axis, int side) {
    const ScalFace* faces[6] = {&f.x_minus, &f.x_plus, &f.y_minus, &f.y_plus, &f.z_minus, &f.z_plus};
    return *faces[axis * 2 + side];
}

float face_velocity(const FlowFace& f, int comp) { return comp == 0 ? f.velocity_x : comp == 1 ? f.velocity_y : f.velocity_z; }

bool scal_periodic(const Scal& s, int a) { return scal_face(s, a, 0).type == kScalarPeriodic && scal_face(s, a, 1).type == kScalarPeriodic; }

Geo make_geo(const SmokeSimulationConfig* cfg) {
    Geo g{};
    g.n[0] = cfg->nx;
    g.n[1] = cfg->ny;
    g.n[2] = cfg->nz;
    g.h = cfg->cell_size;
    g.flow = &cfg->flow_boundary;
    g.density_bc = &cfg->density_boundary;
    g.temperature_bc = &cfg->temperature_boundary;
    for (int a = 0; a < 3; ++a) g.flow_periodic[a] = flow_face(cfg->flow_boundary, a, 0).type == kPeriodic && flow_face(cfg->flow_boundary, a, 1).type == kPeriodic;
    g.cells = int64_t(g.n[0]) * g.n[1] * g.n[2];
    for (int c = 0; c < 3; ++c) g.faces[c] = int64_t(g.n[0] + (c == 0)) * (g.n[1] + (c == 1)) * (g.n[2] + (c == 2));
    return g;
}

// B:121-135 -- x-fastest cell and staggered face indexing.
inline int64_t cidx(const Geo& g, int x, int y, int z) { return (int64_t(z) * g.n[1] + y) * g.n[0] + x; }
inline int64_t fidx(const Geo& g, int c, int i, int j, int k) {
    const int d0 = g.n[0] + (c == 0), d1 = g.n[1] + (c == 1);
    return (int64_t(k) * d1 + j) * d0 + i;
}

// B:137-142
inline int wrap(int v, int n) {
    if (n <= 0) return 0;
    v %= n;
    if (v < 0) v += n;
    return v;
}

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// B:148-160 -- periodic wrap on flow-periodic axes, then bounds test.
bool resolve_flow(const Geo& g, int c[3]) {
    for (int a = 0; a < 3; ++a)
        if (g.flow_periodic[a] && g.n[a] > 0) c[a] = wrap(c[a], g.n[a]);
    return c[0] >= 0 && c[0] < g.n[0] && c[1] >= 0 && c[1] < g.n[1] && c[2] >= 0 && c[2] < g.n[2];
}

bool resolve_scalar(const Geo& g, const Scal& bc, int c[3]) {
    for (int a = 0; a < 3; ++a)
        if (scal_periodic(bc, a) && g.n[a] > 0) c[a] = wrap(c[a], g.n[a]);
    return c[0] >= 0 && c[0] < g.n[0] && c[1] >= 0 && c[1] < g.n[1] && c[2] >= 0 && c[2] < g.n[2];
}

// B:162-166 -- out-of-domain counts as solid.
bool load_occupancy(const Geo& g, const uint8_t* occ, int x, int y, int z) {
    int c[3] = {x, y, z};
    if (!resolve_flow(g, c)) return true;
    return occ[cidx(g, c[0], c[1], c[2])] != 0;
}

// B:168-200 -- scalar ghost: wrap, clamp (zero flux) or the face value; the
// first non-periodic out-of-range axis decides.
template <class R>
R load_scalar(const Geo& g, const Scal& bc, const R* f, int x, int y, int z) {
    int c[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < g.n[a]) continue;
        const ScalFace& face = scal_face(bc, a, c[a] < 0 ? 0 : 1);
        if (scal_periodic(bc, a) && g.n[a] > 0) {
            c[a] = wrap(c[a], g.n[a]);
        } else if (face.type == kZeroFlux && g.n[a] > 0) {
            c[a] = c[a] < 0 ? 0 : g.n[a] - 1;
        } else {
            return R(face.value);
        }
    }
    return f[cidx(g, c[0], c[1], c[2])];
}

// B:202-225 -- wrap or clamp.
template <class R>
R load_flow_cell(const Geo& g, const R* f, int x, int y, int z) {
    int c[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < g.n[a]) continue;
        if (g.flow_periodic[a] && g.n[a] > 0) c[a] = wrap(c[a], g.n[a]);
        else c[a] = c[a] < 0 ? 0 : g.n[a] - 1;
    }
    return f[cidx(g, c[0], c[1], c[2])];
}

// B:227-274 -- cell-centred velocity component with flow ghost rules.
template <class R>
R load_center_component(const Geo& g, const R* f, int comp, int x, int y, int z) {
    int c[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < g.n[a]) continue;
        const FlowFace& face = flow_face(*g.flow, a, c[a] < 0 ? 0 : 1);
        if (g.flow_periodic[a] && g.n[a] > 0) {
            c[a] = wrap(c[a], g.n[a]);
            continue;
        }
        int cl[3];
        for (int b = 0; b < 3; ++b) cl[b] = b == a ? (c[a] < 0 ? 0 : g.n[a] - 1) : clampi(c[b], 0, g.n[b] - 1);
        const R interior = f[cidx(g, cl[0], cl[1], cl[2])];
        const R prescribed = R(face_velocity(face, comp));
        if (face.type == kOutflow) return interior;
        if (face.type == kFreeSlip && comp != a) return interior;
        return R(2) * prescribed - interior;
    }
    return f[cidx(g, c[0], c[1], c[2])];
}

// B:276-376 -- MAC face component `comp` with flow ghost rules.  On the
// component's own (normal) axis the valid range is 0..n, elsewhere 0..n-1.
template <class R>
R load_face(const Geo& g, const R* f, int comp, int i, int j, int k) {
    int c[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) {
        const int hi = g.n[a] - (a == comp ? 0 : 1);
        if (c[a] >= 0 && c[a] <= hi) continue;
        const FlowFace& face = flow_face(*g.flow, a, c[a] < 0 ? 0 : 1);
        if (g.flow_periodic[a] && g.n[a] > 0) {
            c[a] = wrap(c[a], g.n[a]);
            continue;
        }
        int cl[3];
        for (int b = 0; b < 3; ++b) cl[b] = b == a ? (c[a] < 0 ? 0 : hi) : clampi(c[b], 0, g.n[b] - (b == comp ? 0 : 1));
        const R interior = f[fidx(g, comp, cl[0], cl[1], cl[2])];
        if (a == comp) {
            if (face.type == kOutflow) return interior;
        } else if (face.type == kOutflow || face.type == kFreeSlip) {
            return interior;
        }
        return R(2) * R(face_velocity(face, comp)) - interior;
    }
    return f[fidx(g, comp, c[0], c[1], c[2])];
}

// B:378-392 -- Hermite with slope zeroing; result not clamped.
template <class R>
R monotonic_cubic(R p0, R p1, R p2, R p3, R t) {
    const R delta = p2 - p1;
    R m1 = R(0.5) * (p2 - p0);
    R m2 = R(0.5) * (p3 - p1);
    if (std::fabs(delta) < R(1.0e-6)) {
        m1 = R(0);
        m2 = R(0);
    } else {
        if (m1 * delta <= R(0)) m1 = R(0);
        if (m2 * delta <= R(0)) m2 = R(0);
    }
    const R t2 = t * t;
    const R t3 = t2 * t;
    return (R(2) * t3 - R(3) * t2 + R(1)) * p1 + (t3 - R(2) * t2 + t) * m1 + (R(-2) * t3 + R(3) * t2) * p2 + (t3 - t2) * m2;
}

// Field descriptor for the samplers: scalar (with its boundary) or face component.
struct FieldKind {
    int comp;          // -1 scalar, 0..2 face component
    const Scal* bc;    // scalar boundary
};

template <class R>
R load_field(const Geo& g, FieldKind k, const R* f, int i, int j, int l) {
    return k.comp < 0 ? load_scalar(g, *k.bc, f, i, j, l) : load_face(g, f, k.comp, i, j, l);
}

// B:394-754 -- linear (8 taps) and monotone-cubic (64 taps) samplers.  The
// position is wrapped with fmod on periodic axes; grid coordinates are
// p/h - 0.5 except on a face component's own axis (p/h).
template <class R>
R sample(const Geo& g, FieldKind k, bool cubic, const R* f, R x, R y, R z) {
    R p[3] = {x, y, z};
    const R h = R(g.h);
    for (int a = 0; a < 3; ++a) {
        const bool periodic = k.comp < 0 ? scal_periodic(*k.bc, a) : g.flow_periodic[a];
        if (periodic && g.n[a] > 0) {
            const R extent = R(g.n[a]) * h;
            p[a] = std::fmod(p[a], extent);
            if (p[a] < R(0)) p[a] += extent;
        }
    }
    R gc[3];
    int b[3];
    R t[3];
    for (int a = 0; a < 3; ++a) {
        // beyond two cells outside the domain all taps equal the same ghost
        // value: the clamp changes no finite result and keeps indices in range
        gc[a] = std::fmin(std::fmax(a == k.comp ? p[a] / h : p[a] / h - R(0.5), R(-4)), R(g.n[a]) + R(4));
        b[a] = int(std::floor(gc[a]));
        t[a] = gc[a] - R(b[a]);
    }
    if (!cubic) {
        const R c000 = load_field(g, k, f, b[0], b[1], b[2]);
        const R c100 = load_field(g, k, f, b[0] + 1, b[1], b[2]);
        const R c010 = load_field(g, k, f, b[0], b[1] + 1, b[2]);
        const R c110 = load_field(g, k, f, b[0] + 1, b[1] + 1, b[2]);
        const R c001 = load_field(g, k, f, b[0], b[1], b[2] + 1);
        const R c101 = load_field(g, k, f, b[0] + 1, b[1], b[2] + 1);
        const R c011 = load_field(g, k, f, b[0], b[1] + 1, b[2] + 1);
        const R c111 = load_field(g, k, f, b[0] + 1, b[1] + 1, b[2] + 1);
        const R c00 = c000 + (c100 - c000) * t[0];
        const R c10 = c010 + (c110 - c010) * t[0];
        const R c01 = c001 + (c101 - c001) * t[0];
        const R c11 = c011 + (c111 - c011) * t[0];
        const R c0 = c00 + (c10 - c00) * t[1];
        const R c1 = c01 + (c11 - c01) * t[1];
        return c0 + (c1 - c0) * t[2];
    }
    R zs[4];
    for (int dz = 0; dz < 4; ++dz) {
        R ys[4];
        for (int dy = 0; dy < 4; ++dy) {
            const int yy = b[1] + dy - 1, zz = b[2] + dz - 1;
            const R p0 = load_field(g, k, f, b[0] - 1, yy, zz);
            const R p1 = load_field(g, k, f, b[0], yy, zz);
            const R p2 = load_field(g, k, f, b[0] + 1, yy, zz);
            const R p3 = load_field(g, k, f, b[0] + 2, yy, zz);
            ys[dy] = monotonic_cubic(p0, p1, p2, p3, t[0]);
        }
        zs[dz] = monotonic_cubic(ys[0], ys[1], ys[2], ys[3], t[1]);
    }
    return monotonic_cubic(zs[0], zs[1], zs[2], zs[3], t[2]);
}

// Hit test of B:785-794 / 819-828: outside [0, n h] or inside a solid cell
// (n clamps to n-1).  Positions are wrapped on periodic axes first.
template <class R>
bool hits_solid(const Geo& g, const uint8_t* occ, const R q_in[3]) {
    const R h = R(g.h);
    R q[3] = {q_in[0], q_in[1], q_in[2]};
    for (int a = 0; a < 3; ++a) {
        if (g.flow_periodic[a] && g.n[a] > 0) {
            const R extent = R(g.n[a]) * h;
            q[a] = std::fmod(q[a], extent);
            if (q[a] < R(0)) q[a] += extent;
        }
    }
    for (int a = 0; a < 3; ++a)
        if (q[a] < R(0) || q[a] > R(g.n[a]) * h) return true;
    int c[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = int(std::floor(q[a] / h));
        if (c[a] == g.n[a]) c[a] = g.n[a] - 1;
    }
    if (c[0] < 0 || c[0] >= g.n[0] || c[1] < 0 || c[1] >= g.n[1] || c[2] < 0 || c[2] >= g.n[2]) return true;
    return occ[cidx(g, c[0], c[1], c[2])] != 0;
}

// B:756-838 -- midpoint backtrace with linearly sampled velocity; a clipped
// path returns start + (end - start) * lo after 10 bisection steps.  The
// returned point is not wrapped.
template <class R>
void trace_rk2(const Geo& g, const R* u, const R* v, const R* w, const uint8_t* occ, R dt, const R start[3], R out[3]) {
    const FieldKind ku{0, nullptr}, kv{1, nullptr}, kw{2, nullptr};
    const R v0[3] = {sample(g, ku, false, u, start[0], start[1], start[2]), sample(g, kv, false, v, start[0], start[1], start[2]), sample(g, kw, false, w, start[0], start[1], start[2])};
    const R mid[3] = {start[0] - R(0.5) * dt * v0[0], start[1] - R(0.5) * dt * v0[1], start[2] - R(0.5) * dt * v0[2]};
    const R v1[3] = {sample(g, ku, false, u, mid[0], mid[1], mid[2]), sample(g, kv, false, v, mid[0], mid[1], mid[2]), sample(g, kw, false, w, mid[0], mid[1], mid[2])};
    const R traced[3] = {start[0] - dt * v1[0], start[1] - dt * v1[1], start[2] - dt * v1[2]};
    if (!hits_solid(g, occ, traced)) {
        out[0] = traced[0];
        out[1] = traced[1];
        out[2] = traced[2];
        return;
    }
    R lo = R(0), hi = R(1);
    for (int it = 0; it < 10; ++it) {
        const R mt = R(0.5) * (lo + hi);
        const R test[3] = {start[0] + (traced[0] - start[0]) * mt, start[1] + (traced[1] - start[1]) * mt, start[2] + (traced[2] - start[2]) * mt};
        if (hits_solid(g, occ, test)) hi = mt;
        else lo = mt;
    }
    for (int a = 0; a < 3; ++a) out[a] = start[a] + (traced[a] - start[a]) * lo;
}

// B:1007-1012 / 1288-1305 -- mean solid velocity of the solid cells beside a
// face; `has_override` false when neither side is solid.
template <class R>
bool solid_face_override(const Geo& g, const uint8_t* occ, const R* solid_vel, int comp, int i, int j, int k, R& value) {
    int left[3] = {i, j, k};
    int right[3] = {i, j, k};
    left[comp] -= 1;
    const bool has_left = resolve_flow(g, left);
    const bool has_right = resolve_flow(g, right);
    const bool left_occ = has_left && occ[cidx(g, left[0], left[1], left[2])] != 0;
    const bool right_occ = has_right && occ[cidx(g, right[0], right[1], right[2])] != 0;
    if (!left_occ && !right_occ) return false;
    R sum = R(0), weight = R(0);
    if (left_occ) {
        sum += solid_vel[cidx(g, left[0], left[1], left[2])];
        weight += R(1);
    }
    if (right_occ) {
        sum += solid_vel[cidx(g, right[0], right[1], right[2])];
        weight += R(1);
    }
    value = weight > R(0) ? sum / weight : R(0);
    return true;
}

bool is_domain_face(const Geo& g, int comp, int idx_on_axis) {
    if (idx_on_axis == 0) return flow_face(*g.flow, comp, 0).type != kPeriodic;
    if (idx_on_axis == g.n[comp]) return flow_face(*g.flow, comp, 1).type != kPeriodic;
    return false;
}

// Iterate every face of component `comp`.
template <class F>
void for_each_face(const Geo& g, int comp, F&& fn) {
    const int d[3] = {g.n[0] + (comp == 0), g.n[1] + (comp == 1), g.n[2] + (comp == 2)};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) fn(i, j, k);
}

template <class F>
void for_each_cell(const Geo& g, F&& fn) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g.n[2]; ++z)
        for (int y = 0; y < g.n[1]; ++y)
            for (int x = 0; x < g.n[0]; ++x) fn(x, y, z);
}

// B:1014-1189 -- domain-face and solid-face velocity, then periodic sync.
// Interior faces first; domain faces then read the finished interior values
// (both ends of a line are read before either is written).
template <class R>
void enforce_component(const Geo& g, R* f, int comp, const uint8_t* occ, const R* solid_vel) {
    for_each_face(g, comp, [&](int i, int j, int k) {
        const int c[3] = {i, j, k};
        if (is_domain_face(g, comp, c[comp])) return;
        R value;
        if (solid_face_override(g, occ, solid_vel, comp, i, j, k, value)) f[fidx(g, comp, i, j, k)] = value;
    });
    const int a1 = (comp + 1) % 3, a2 = (comp + 2) % 3;
    const int n = g.n[comp];
    const FlowFace& lo = flow_face(*g.flow, comp, 0);
    const FlowFace& hi = flow_face(*g.flow, comp, 1);
    for (int q2 = 0; q2 < g.n[a2]; ++q2) {
        for (int q1 = 0; q1 < g.n[a1]; ++q1) {
            auto at = [&](int idx) -> R& {
                int c[3];
                c[comp] = idx;
                c[a1] = q1;
                c[a2] = q2;
                return f[fidx(g, comp, c[0], c[1], c[2])];
            };
            const R inner_lo = at(1);
            const R inner_hi = at(n - 1);
            if (lo.type != kPeriodic) {
                R& face = at(0);
                face = lo.type == kOutflow ? inner_lo : R(face_velocity(lo, comp));
            }
            if (hi.type != kPeriodic) {
                R& face = at(n);
                face = hi.type == kOutflow ? inner_hi : R(face_velocity(hi, comp));
            }
            if (g.flow_periodic[comp]) at(n) = at(0);
        }
    }
}

template <class R>
void enforce_all(const Geo& g, R* u, R* v, R* w, const uint8_t* occ, const R* svx, const R* svy, const R* svz) {
    enforce_component(g, u, 0, occ, svx);
    enforce_component(g, v, 1, occ, svy);
    enforce_component(g, w, 2, occ, svz);
}

// B:859-868
template <class R>
void center_velocity(const Geo& g, const R* u, const R* v, const R* w, R* cx, R* cy, R* cz) {
    for_each_cell(g, [&](int x, int y, int z) {
        const int64_t i = cidx(g, x, y, z);
        cx[i] = R(0.5) * (u[fidx(g, 0, x, y, z)] + u[fidx(g, 0, x + 1, y, z)]);
        cy[i] = R(0.5) * (v[fidx(g, 1, x, y, z)] + v[fidx(g, 1, x, y + 1, z)]);
        cz[i] = R(0.5) * (w[fidx(g, 2, x, y, z)] + w[fidx(g, 2, x, y, z + 1)]);
    });
}

// B:870-900
template <class R>
void vorticity(const Geo& g, const R* cx, const R* cy, const R* cz, const uint8_t* occ, R* ox, R* oy, R* oz, R* om) {
    const R h = R(g.h);
    for_each_cell(g, [&](int x, int y, int z) {
        const int64_t i = cidx(g, x, y, z);
        if (load_occupancy(g, occ, x, y, z)) {
            ox[i] = oy[i] = oz[i] = om[i] = R(0);
            return;
        }
        const R dvz_dy = R(0.5) * (load_center_component(g, cz, 2, x, y + 1, z) - load_center_component(g, cz, 2, x, y - 1, z)) / h;
        const R dvy_dz = R(0.5) * (load_center_component(g, cy, 1, x, y, z + 1) - load_center_component(g, cy, 1, x, y, z - 1)) / h;
        const R dvx_dz = R(0.5) * (load_center_component(g, cx, 0, x, y, z + 1) - load_center_component(g, cx, 0, x, y, z - 1)) / h;
        const R dvz_dx = R(0.5) * (load_center_component(g, cz, 2, x + 1, y, z) - load_center_component(g, cz, 2, x - 1, y, z)) / h;
        const R dvy_dx = R(0.5) * (load_center_component(g, cy, 1, x + 1, y, z) - load_center_component(g, cy, 1, x - 1, y, z)) / h;
        const R dvx_dy = R(0.5) * (load_center_component(g, cx, 0, x, y + 1, z) - load_center_component(g, cx, 0, x, y - 1, z)) / h;
        const R wx = dvz_dy - dvy_dz, wy = dvx_dz - dvz_dx, wz = dvy_dx - dvx_dy;
        ox[i] = wx;
        oy[i] = wy;
        oz[i] = wz;
        om[i] = std::sqrt(wx * wx + wy * wy + wz * wz);
    });
}

// B:902-948 -- user force, buoyancy (+y), vorticity confinement.
template <class R>
void cell_forces(const Geo& g, const SmokeSimulationConfig* cfg, const R* fux, const R* fuy, const R* fuz, const R* rho, const R* T, const uint8_t* occ,
                 const R* ox, const R* oy, const R* oz, const R* om, R* fx, R* fy, R* fz) {
    const R h = R(g.h);
    const R a = R(cfg->buoyancy_density_factor), b = R(cfg->buoyancy_temperature_factor), tamb = R(cfg->ambient_temperature);
    const R eps = R(cfg->vorticity_confinement);
    for_each_cell(g, [&](int x, int y, int z) {
        const int64_t i = cidx(g, x, y, z);
        fx[i] = fux ? fux[i] : R(0);
        fy[i] = fuy ? fuy[i] : R(0);
        fz[i] = fuz ? fuz[i] : R(0);
        if (load_occupancy(g, occ, x, y, z)) return;
        fy[i] += -a * rho[i] + b * (T[i] - tamb);
        const R gx = R(0.5) * (load_flow_cell(g, om, x + 1, y, z) - load_flow_cell(g, om, x - 1, y, z)) / h;
        const R gy = R(0.5) * (load_flow_cell(g, om, x, y + 1, z) - load_flow_cell(g, om, x, y - 1, z)) / h;
        const R gz = R(0.5) * (load_flow_cell(g, om, x, y, z + 1) - load_flow_cell(g, om, x, y, z - 1)) / h;
        const R gm = std::sqrt(gx * gx + gy * gy + gz * gz);
        if (gm < R(1.0e-6)) return;
        const R inv = R(1) / gm;
        const R nx_ = gx * inv, ny_ = gy * inv, nz_ = gz * inv;
        fx[i] += eps * h * (ny_ * oz[i] - nz_ * oy[i]);
        fy[i] += eps * h * (nz_ * ox[i] - nx_ * oz[i]);
        fz[i] += eps * h * (nx_ * oy[i] - ny_ * ox[i]);
    });
}

// B:950-1005 -- average of the in-domain cell forces beside each face (no wrap).
template <class R>
void add_forces_to_faces(const Geo& g, R* f, int comp, const R* force, R dt) {
    for_each_face(g, comp, [&](int i, int j, int k) {
        const int c[3] = {i, j, k};
        R sum = R(0), weight = R(0);
        if (c[comp] > 0) {
            int l[3] = {i, j, k};
            l[comp] -= 1;
            sum += force[cidx(g, l[0], l[1], l[2])];
            weight += R(1);
        }
        if (c[comp] < g.n[comp]) {
            sum += force[cidx(g, i, j, k)];
            weight += R(1);
        }
        if (weight > R(0)) f[fidx(g, comp, i, j, k)] += dt * (sum / weight);
    });
}

// B:1191-1219 -- semi-Lagrangian advection of every face of one component.
template <class R>
void advect_component(const Geo& g, const R* u, const R* v, const R* w, const R* src, R* dst, int comp, const uint8_t* occ, R dt, bool cubic) {
    const R h = R(g.h);
    const FieldKind k{comp, nullptr};
    for_each_face(g, comp, [&](int i, int j, int l) {
        const int c[3] = {i, j, l};
        R start[3];
        for (int a = 0; a < 3; ++a) start[a] = a == comp ? R(c[a]) * h : (R(c[a]) + R(0.5)) * h;
        R p[3];
        trace_rk2(g, u, v, w, occ, dt, start, p);
        dst[fidx(g, comp, i, j, l)] = sample(g, k, cubic, src, p[0], p[1], p[2]);
    });
}

// B:1221-1233 -- scalar advection, 0 in solid cells.
template <class R>
void advect_scalar(const Geo& g, const Scal& bc, const R* u, const R* v, const R* w, const R* src, R* dst, const uint8_t* occ, R dt, bool cubic) {
    const R h = R(g.h);
    const FieldKind k{-1, &bc};
    for_each_cell(g, [&](int x, int y, int z) {
        const int64_t i = cidx(g, x, y, z);
        if (load_occupancy(g, occ, x, y, z)) {
            dst[i] = R(0);
            return;
        }
        const R start[3] = {(R(x) + R(0.5)) * h, (R(y) + R(0.5)) * h, (R(z) + R(0.5)) * h};
        R p[3];
        trace_rk2(g, u, v, w, occ, dt, start, p);
        dst[i] = sample(g, k, cubic, src, p[0], p[1], p[2]);
    });
}

// B:1243-1290 -- nearest fluid cell by smallest Chebyshev shell, then least
// squared distance, first in dz,dy,dx order on ties; -1 if none.
int64_t nearest_fluid(const Geo& g, const Scal& bc, const uint8_t* occ, int x, int y, int z) {
    const int max_radius = std::max(g.n[0], std::max(g.n[1], g.n[2]));
    for (int r = 1; r <= max_radius; ++r) {
        bool found = false;
        int best_d2 = 0;
        int64_t best = -1;
        for (int dz = -r; dz <= r; ++dz)
            for (int dy = -r; dy <= r; ++dy)
                for (int dx = -r; dx <= r; ++dx) {
                    const int shell = std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz)));
                    if (shell != r) continue;
                    int c[3] = {x + dx, y + dy, z + dz};
                    if (!resolve_scalar(g, bc, c)) continue;
                    const int64_t ni = cidx(g, c[0], c[1], c[2]);
                    if (occ[ni] != 0) continue;
                    const int d2 = dx * dx + dy * dy + dz * dz;
                    if (!found || d2 < best_d2) {
                        found = true;
                        best_d2 = d2;
                        best = ni;
                    }
                }
        if (found) return best;
    }
    return -1;
}

// B:1292-1302 -- first fluid cell; cell count if there is none.
int64_t pressure_anchor(const Geo& g, const uint8_t* occ) {
    for (int64_t i = 0; i < g.cells; ++i)
        if (occ[i] == 0) return i;
    return g.cells;
}

// Rows of the Poisson matrix as built by B:1343-1412 on the pattern of
// B:1688-1752: identity rows for solid cells and the anchor; otherwise the
// diagonal counts fluid neighbours (anchor included) and OUTFLOW domain faces,
// off-diagonals are -1 for distinct fluid non-anchor neighbours (a neighbour
// equal to the row itself only adds to the diagonal).
struct PoissonRows {
    std::vector<double> diag;
    std::vector<uint8_t> identity;
    std::vector<int64_t> cols;   // 6 per row, -1 unused
    int64_t anchor = 0;
};

PoissonRows build_rows(const Geo& g, const uint8_t* occ) {
    PoissonRows rows;
    rows.anchor = pressure_anchor(g, occ);
    rows.diag.assign(size_t(g.cells), 1.0);
    rows.identity.assign(size_t(g.cells), 1);
    rows.cols.assign(size_t(g.cells) * 6, -1);
    for_each_cell(g, [&](int x, int y, int z) {
        const int64_t row = cidx(g, x, y, z);
        if (occ[row] != 0 || row == rows.anchor) return;
        rows.identity[size_t(row)] = 0;
        double diag = 0.0;
        int count = 0;
        int64_t* out = &rows.cols[size_t(row) * 6];
        const int base[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) {
            for (int s = -1; s <= 1; s += 2) {
                int c[3] = {base[0], base[1], base[2]};
                c[a] += s;
                if (c[a] < 0 || c[a] >= g.n[a]) {
                    if (g.flow_periodic[a]) {
                        c[a] = wrap(c[a], g.n[a]);
                    } else {
                        if (flow_face(*g.flow, a, s < 0 ? 0 : 1).type == kOutflow) diag += 1.0;
                        continue;
                    }
                }
                const int64_t nb = cidx(g, c[0], c[1], c[2]);
                if (occ[nb] != 0) continue;
                diag += 1.0;
                if (nb == rows.anchor || nb == row) continue;
                bool dup = false;
                for (int q = 0; q < count; ++q) dup = dup || out[q] == nb;
                if (!dup) out[count++] = nb;
            }
        }
        rows.diag[size_t(row)] = diag <= 0.0 ? 1.0 : diag;
    });
    return rows;
}

template <class R, class Acc>
void apply_rows(const Geo& g, const PoissonRows& rows, const R* x, Acc* y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < g.cells; ++i) {
        Acc s = Acc(rows.diag[size_t(i)]) * Acc(x[i]);
        const int64_t* c = &rows.cols[size_t(i) * 6];
        for (int q = 0; q < 6 && c[q] >= 0; ++q) s -= Acc(x[c[q]]);
        y[i] = s;
    }
}

// Deterministic FP64 dot: per-plane partials, summed in plane order.
template <class A, class B>
double dot(const Geo& g, const A* a, const B* b) {
    const int64_t plane = int64_t(g.n[0]) * g.n[1];
    std::vector<double> part(size_t(g.n[2]), 0.0);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g.n[2]; ++z) {
        double s = 0.0;
        for (int64_t i = z * plane; i < (z + 1) * plane; ++i) s += double(a[i]) * double(b[i]);
        part[size_t(z)] = s;
    }
    double s = 0.0;
    for (double v : part) s += v;
    return s;
}

// B:1304-1328 -- b = -(h^2/dt) div(u*) + sum of OUTFLOW face pressures; 0 at
// the anchor and in solids.
template <class R>
void pressure_rhs(const Geo& g, const uint8_t* occ, int64_t anchor, const R* us, const R* vs, const R* ws, R dt, R* rhs) {
    const R h = R(g.h);
    const Flow& fb = *g.flow;
    for_each_cell(g, [&](int x, int y, int z) {
        const int64_t i = cidx(g, x, y, z);
        if (i == anchor || occ[i] != 0) {
            rhs[i] = R(0);
            return;
        }
        const R div = (us[fidx(g, 0, x + 1, y, z)] - us[fidx(g, 0, x, y, z)] + vs[fidx(g, 1, x, y + 1, z)] - vs[fidx(g, 1, x, y, z)] + ws[fidx(g, 2, x, y, z + 1)] - ws[fidx(g, 2, x, y, z)]) / h;
        R bsum = R(0);
        if (x == 0 && fb.x_minus.type == kOutflow) bsum += R(fb.x_minus.pressure);
        if (x == g.n[0] - 1 && fb.x_plus.type == kOutflow) bsum += R(fb.x_plus.pressure);
        if (y == 0 && fb.y_minus.type == kOutflow) bsum += R(fb.y_minus.pressure);
        if (y == g.n[1] - 1 && fb.y_plus.type == kOutflow) bsum += R(fb.y_plus.pressure);
        if (z == 0 && fb.z_minus.type == kOutflow) bsum += R(fb.z_minus.pressure);
        if (z == g.n[2] - 1 && fb.z_plus.type == kOutflow) bsum += R(fb.z_plus.pressure);
        rhs[i] = -(h * h / dt) * div + bsum;
    });
}

// B:1414-1418 guards the CG ratios with |den| > 1e-20 (fixed-K mode keeps it);
// the tolerance solve only guards against an exact zero, since small right-hand
// sides legitimately drive r.z below 1e-20 before the target is reached.
inline double guarded_ratio(double num, double den, double limit = 1.0e-20) { return std::fabs(den) > limit ? num / den : 0.0; }

// B:1800-1823 -- fixed-K unpreconditioned CG from x = 0 (dots in FP64).
template <class R>
void solve_fixed(const Geo& g, const PoissonRows& rows, const R* b, double* x_out, int iterations, OracleStats* st) {
    const size_t n = size_t(g.cells);
    std::vector<R> x(n, R(0)), r(b, b + n), p(b, b + n), ap(n);
    double rz = dot(g, r.data(), r.data());
    for (int it = 0; it < iterations; ++it) {
        apply_rows(g, rows, p.data(), ap.data());
        const double pap = dot(g, p.data(), ap.data());
        const R alpha = R(guarded_ratio(rz, pap));
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < int64_t(n); ++i) {
            x[size_t(i)] += alpha * p[size_t(i)];
            r[size_t(i)] -= alpha * ap[size_t(i)];
        }
        const double rr = dot(g, r.data(), r.data());
        const R beta = R(guarded_ratio(rr, rz));
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < int64_t(n); ++i) p[size_t(i)] = beta * p[size_t(i)] + r[size_t(i)];
        rz = rr;
    }
    for (size_t i = 0; i < n; ++i) x_out[i] = double(x[i]);
    st->iterations = iterations;
    st->refinements = 0;
    std::vector<double> ax(n);
    apply_rows(g, rows, x.data(), ax.data());
    double rr = 0.0, bb = 0.0;
    for (size_t i = 0; i < n; ++i) {
        rr += (double(b[i]) - ax[i]) * (double(b[i]) - ax[i]);
        bb += double(b[i]) * double(b[i]);
    }
    st->rhs_norm = std::sqrt(bb);
    st->relative_residual = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
    st->converged = 0;
}

// Tolerance-terminated PCG with FP64 outer refinement: x (FP64) is refined by
// inner solves A e = r carried in R with FP64 reductions; the true residual
// b - A x is evaluated in FP64 after every inner solve.
template <class R>
void solve_tolerance(const Geo& g, const PoissonRows& rows, const OracleOptions* opt, const R* b, double* x, OracleStats* st) {
    const size_t n = size_t(g.cells);
    const bool jacobi = opt->preconditioner == 1;
    if (!opt->warm_start) std::fill(x, x + n, 0.0);
    for (size_t i = 0; i < n; ++i)
        if (rows.identity[i]) x[i] = 0.0;
    std::vector<double> r64(n), ax(n);
    std::vector<R> r(n), e(n), p(n), ap(n);
    const double bnorm = std::sqrt(dot(g, b, b));
    auto true_residual = [&]() {
        apply_rows(g, rows, x, ax.data());
        for (size_t i = 0; i < n; ++i) r64[i] = double(b[i]) - ax[i];
        return std::sqrt(dot(g, r64.data(), r64.data()));
    };
    double rnorm = true_residual();
    int iterations = 0, passes = 0;
    const double target_abs = opt->relative_tolerance * bnorm;
    while (rnorm > target_abs && passes <= opt->max_refinements && iterations < opt->max_iterations) {
        for (size_t i = 0; i < n; ++i) {
            r[i] = R(r64[i]);
            e[i] = R(0);
            p[i] = R(0);
        }
        const double inner_target = std::max(0.5 * target_abs, opt->inner_reduction * rnorm);
        double beta = 0.0, rz = 1.0;
        bool first = true;
        while (iterations < opt->max_iterations) {
            // p = z + beta p, z = M^-1 r
            double rz_new = 0.0;
            {
                std::vector<double> part(size_t(g.n[2]), 0.0);
                const int64_t plane = int64_t(g.n[0]) * g.n[1];
#pragma omp parallel for schedule(static)
                for (int z = 0; z < g.n[2]; ++z) {
                    double s = 0.0;
                    for (int64_t i = z * plane; i < (z + 1) * plane; ++i) {
                        const R zi = jacobi ? R(r[size_t(i)] / R(rows.diag[size_t(i)])) : r[size_t(i)];
                        s += double(r[size_t(i)]) * double(zi);
                    }
                    part[size_t(z)] = s;
                }
                for (double v : part) rz_new += v;
            }
            if (!first) beta = guarded_ratio(rz_new, rz, 1.0e-300);
            rz = rz_new;
            first = false;
            const R rbeta = R(beta);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < int64_t(n); ++i) {
                const R zi = jacobi ? R(r[size_t(i)] / R(rows.diag[size_t(i)])) : r[size_t(i)];
                p[size_t(i)] = zi + rbeta * p[size_t(i)];
            }
            apply_rows(g, rows, p.data(), ap.data());
            const double pap = dot(g, p.data(), ap.data());
            const R alpha = R(guarded_ratio(rz, pap, 1.0e-300));
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < int64_t(n); ++i) {
                e[size_t(i)] += alpha * p[size_t(i)];
                r[size_t(i)] -= alpha * ap[size_t(i)];
            }
            ++iterations;
            const double rr = dot(g, r.data(), r.data());
            if (std::sqrt(rr) <= inner_target) break;
        }
        for (size_t i = 0; i < n; ++i) x[i] += double(e[i]);
        rnorm = true_residual();
        ++passes;
    }
    st->iterations = iterations;
    st->refinements = passes > 0 ? passes - 1 : 0;
    st->rhs_norm = bnorm;
    st->relative_residual = bnorm > 0.0 ? rnorm / bnorm : 0.0;
    st->converged = rnorm <= target_abs ? 1 : 0;
}

// B:1425-1600 -- u* -= dt (p_R - p_L)/h on interior faces, solid faces take
// the solid velocity; domain faces are rewritten by the following enforce.
template <class R>
void project_component(const Geo& g, R* f, int comp, const R* p, const uint8_t* occ, const R* solid_vel, R dt) {
    const R h = R(g.h);
    for_each_face(g, comp, [&](int i, int j, int k) {
        const int c[3] = {i, j, k};
        if (is_domain_face(g, comp, c[comp])) return;
        R value;
        if (solid_face_override(g, occ, solid_vel, comp, i, j, k, value)) {
            f[fidx(g, comp, i, j, k)] = value;
            return;
        }
        int left[3] = {i, j, k};
        int right[3] = {i, j, k};
        left[comp] -= 1;
        const bool hl = resolve_flow(g, left), hr = resolve_flow(g, right);
        if (hl && hr) {
            const R pr = p[cidx(g, right[0], right[1], right[2])];
            const R pl = p[cidx(g, left[0], left[1], left[2])];
            f[fidx(g, comp, i, j, k)] -= dt * (pr - pl) / h;
        }
    });
}

template <class R>
struct Fields {
    R *u, *v, *w, *us, *vs, *ws, *rho, *T, *rho_src, *T_src, *fux, *fuy, *fuz, *svx, *svy, *svz, *T_solid, *rhs;
    R *fx, *fy, *fz, *cx, *cy, *cz, *ox, *oy, *oz, *om, *div, *vmag;
    double* pressure;
    uint8_t* occ;
};

template <class R>
double max_face_speed(const Geo& g, const Fields<R>& s) {
    double m = 0.0;
    const R* arrays[3] = {s.u, s.v, s.w};
    for (int c = 0; c < 3; ++c)
        for (int64_t i = 0; i < g.faces[c]; ++i) m = std::max(m, double(std::fabs(arrays[c][i])));
    return m;
}

// One step, stages first..last, in the order captured at B:2000-2057.
template <class R>
void run_stages(const SmokeSimulationConfig* cfg, const OracleOptions* opt, Fields<R>& s, int first, int last, OracleStats* st, R& dt_state) {
    const Geo g = make_geo(cfg);
    const bool cubic = cfg->scalar_advection_mode == SMOKE_SIMULATION_SCALAR_ADVECTION_MONOTONIC_CUBIC;
    const size_t n = size_t(g.cells);
    if (first <= 0 && last >= 0) {
        const double speed = max_face_speed(g, s);
        st->max_face_speed = speed;
        if (opt->adaptive_dt) {
            const double cand = speed > 0.0 ? opt->cfl * g.h / speed : opt->dt_max;
            dt_state = R(std::min(opt->dt_max, cand));
        } else {
            dt_state = R(cfg->dt);
        }
        st->dt = double(dt_state);
    }
    const R dt = dt_state;
    if (first <= 0 && last >= 0) {
        // B:2001 solid temperature, B:2002-2003 vorticity, B:2005-2008 forces
        for (size_t i = 0; i < n; ++i)
            if (s.occ[i]) s.T[i] = s.T_solid[i];
        center_velocity(g, s.u, s.v, s.w, s.cx, s.cy, s.cz);
        vorticity(g, s.cx, s.cy, s.cz, s.occ, s.ox, s.oy, s.oz, s.om);
        cell_forces(g, cfg, s.fux, s.fuy, s.fuz, s.rho, s.T, s.occ, s.ox, s.oy, s.oz, s.om, s.fx, s.fy, s.fz);
        // B:2010-2018
        add_forces_to_faces(g, s.u, 0, s.fx, dt);
        add_forces_to_faces(g, s.v, 1, s.fy, dt);
        add_forces_to_faces(g, s.w, 2, s.fz, dt);
        enforce_all(g, s.u, s.v, s.w, s.occ, s.svx, s.svy, s.svz);
    }
    if (first <= 1 && last >= 1) {
        // B:2019-2030
        advect_component(g, s.u, s.v, s.w, s.u, s.us, 0, s.occ, dt, cubic);
        advect_component(g, s.u, s.v, s.w, s.v, s.vs, 1, s.occ, dt, cubic);
        advect_component(g, s.u, s.v, s.w, s.w, s.ws, 2, s.occ, dt, cubic);
        enforce_all(g, s.us, s.vs, s.ws, s.occ, s.svx, s.svy, s.svz);
    }
    if (first <= 2 && last >= 2) {
        // B:2031 -> B:1788-1824
        const PoissonRows rows = build_rows(g, s.occ);
        st->anchor = int32_t(rows.anchor);
        pressure_rhs(g, s.occ, rows.anchor, s.us, s.vs, s.ws, dt, s.rhs);
        if (opt->pressure_mode == 0) solve_fixed(g, rows, s.rhs, s.pressure, cfg->pressure_iterations, st);
        else solve_tolerance(g, rows, opt, s.rhs, s.pressure, st);
    }
    if (first <= 3 && last >= 3) {
        // B:2032-2043; the FP32 state keeps the rounded pressure
        std::vector<R> p(n);
        for (size_t i = 0; i < n; ++i) p[i] = R(s.pressure[i]);
        project_component(g, s.us, 0, p.data(), s.occ, s.svx, dt);
        project_component(g, s.vs, 1, p.data(), s.occ, s.svy, dt);
        project_component(g, s.ws, 2, p.data(), s.occ, s.svz, dt);
        enforce_all(g, s.us, s.vs, s.ws, s.occ, s.svx, s.svy, s.svz);
        std::copy(s.us, s.us + g.faces[0], s.u);
        std::copy(s.vs, s.vs + g.faces[1], s.v);
        std::copy(s.ws, s.ws + g.faces[2], s.w);
    }
    if (first <= 4 && last >= 4) {
        // B:2044-2052
        std::vector<R> tmp(n), adv(n);
        for (size_t i = 0; i < n; ++i) tmp[i] = s.T[i] + dt * s.T_src[i];
        advect_scalar(g, *g.temperature_bc, s.u, s.v, s.w, tmp.data(), s.T, s.occ, dt, cubic);
        for (size_t i = 0; i < n; ++i)
            if (s.occ[i]) s.T[i] = s.T_solid[i];
        for (size_t i = 0; i < n; ++i) tmp[i] = s.rho[i] + dt * s.rho_src[i];
        advect_scalar(g, *g.density_bc, s.u, s.v, s.w, tmp.data(), adv.data(), s.occ, dt, cubic);
        for_each_cell(g, [&](int x, int y, int z) {
            const int64_t i = cidx(g, x, y, z);
            if (s.occ[i] == 0) {
                s.rho[i] = adv[size_t(i)];
                return;
            }
            const int64_t nb = nearest_fluid(g, *g.density_bc, s.occ, x, y, z);
            s.rho[i] = nb >= 0 ? adv[size_t(nb)] : R(0);
        });
    }
}

template <class R>
void diagnostics(const SmokeSimulationConfig* cfg, Fields<R>& s) {
    const Geo g = make_geo(cfg);
    const R h = R(g.h);
    center_velocity(g, s.u, s.v, s.w, s.cx, s.cy, s.cz);
    vorticity(g, s.cx, s.cy, s.cz, s.occ, s.ox, s.oy, s.oz, s.om);
    // B:1330-1341, 1602-1609
    for_each_cell(g, [&](int x, int y, int z) {
        const int64_t i = cidx(g, x, y, z);
        s.div[i] = s.occ[i] ? R(0)
                            : (s.u[fidx(g, 0, x + 1, y, z)] - s.u[fidx(g, 0, x, y, z)] + s.v[fidx(g, 1, x, y + 1, z)] - s.v[fidx(g, 1, x, y, z)] + s.w[fidx(g, 2, x, y, z + 1)] - s.w[fidx(g, 2, x, y, z)]) / h;
        s.vmag[i] = std::sqrt(s.cx[i] * s.cx[i] + s.cy[i] * s.cy[i] + s.cz[i] * s.cz[i]);
    });
}

Fields<double> wrap_state(OracleState* s) {
    Fields<double> f{};
    f.u = s->u; f.v = s->v; f.w = s->w;
    f.us = s->us; f.vs = s->vs; f.ws = s->ws;
    f.rho = s->density; f.T = s->temperature;
    f.rho_src = s->density_source; f.T_src = s->temperature_source;
    f.fux = s->force_user_x; f.fuy = s->force_user_y; f.fuz = s->force_user_z;
    f.svx = s->solid_velocity_x; f.svy = s->solid_velocity_y; f.svz = s->solid_velocity_z;
    f.T_solid = s->solid_temperature; f.rhs = s->pressure_rhs;
    f.fx = s->force_x; f.fy = s->force_y; f.fz = s->force_z;
    f.cx = s->center_x; f.cy = s->center_y; f.cz = s->center_z;
    f.ox = s->vorticity_x; f.oy = s->vorticity_y; f.oz = s->vorticity_z; f.om = s->vorticity_magnitude;
    f.div = s->divergence; f.vmag = s->velocity_magnitude;
    f.pressure = s->pressure;
    f.occ = s->occupancy;
    return f;
}

// Scenario noise: a 32-bit integer mix (the "lowbias32" finaliser published
// by C. Wellons) of (seed, step, cell); the GPU harness uses the same bits.
inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

}  // namespace

extern "C" {

uint32_t oracle_noise_bits(uint32_t seed, uint32_t step, uint32_t cell) { return mix32(cell ^ mix32(step ^ mix32(seed))); }

void oracle_plume_sources(const SmokeSimulationConfig* cfg, const OraclePlumeSource* src, uint32_t step, double* density_source, double* temperature_source) {
    const Geo g = make_geo(cfg);
    for (int z = 0; z < g.n[2]; ++z)
        for (int y = 0; y < g.n[1]; ++y)
            for (int x = 0; x < g.n[0]; ++x) {
                const int64_t i = cidx(g, x, y, z);
                const float dx = (float(x) + 0.5f) - src->center_x;
                const float dy = (float(y) + 0.5f) - src->center_y;
                const float dz = (float(z) + 0.5f) - src->center_z;
                const float r2 = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
                if (r2 > src->radius * src->radius) {
                    density_source[i] = 0.0;
                    temperature_source[i] = 0.0;
                    continue;
                }
                const uint32_t bits = oracle_noise_bits(src->seed, step, uint32_t(i));
                const float unit = float(bits >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
                density_source[i] = double(src->density_rate);
                temperature_source[i] = double(std::fma(src->temperature_noise, unit, src->temperature_rate));
            }
}

int oracle_run_stages(const SmokeSimulationConfig* cfg, const OracleOptions* opt, OracleState* s, int first, int last, OracleStats* stats) {
    // `stats` is in/out: a run that starts after stage 0 takes dt from it.
    Fields<double> f = wrap_state(s);
    OracleStats local{};
    if (stats) local = *stats;
    double dt = local.dt > 0.0 ? local.dt : double(cfg->dt);
    run_stages(cfg, opt, f, first, last, &local, dt);
    local.dt = dt;
    if (stats) *stats = local;
    return 0;
}

void oracle_diagnostics(const SmokeSimulationConfig* cfg, OracleState* s) {
    Fields<double> f = wrap_state(s);
    diagnostics(cfg, f);
}

double oracle_monotonic_cubic(double p0, double p1, double p2, double p3, double t) { return monotonic_cubic(p0, p1, p2, p3, t); }

double oracle_sample(const SmokeSimulationConfig* cfg, int field_kind, int cubic, const double* field, double x, double y, double z) {
    const Geo g = make_geo(cfg);
    FieldKind k{};
    if (field_kind == 0) k = {-1, g.density_bc};
    else if (field_kind == 1) k = {-1, g.temperature_bc};
    else k = {field_kind - 2, nullptr};
    return sample(g, k, cubic != 0, field, x, y, z);
}

double oracle_load(const SmokeSimulationConfig* cfg, int field_kind, const double* field, int i, int j, int k) {
    const Geo g = make_geo(cfg);
    if (field_kind == 0) return load_scalar(g, *g.density_bc, field, i, j, k);
    if (field_kind == 1) return load_scalar(g, *g.temperature_bc, field, i, j, k);
    if (field_kind == 5) return load_flow_cell(g, field, i, j, k);
    return load_face(g, field, field_kind - 2, i, j, k);
}

double oracle_load_center_component(const SmokeSimulationConfig* cfg, int component, const double* field, int x, int y, int z) {
    const Geo g = make_geo(cfg);
    return load_center_component(g, field, component, x, y, z);
}

void oracle_trace(const SmokeSimulationConfig* cfg, const double* u, const double* v, const double* w, const uint8_t* occ, double dt, const double* start, double* out) {
    const Geo g = make_geo(cfg);
    trace_rk2(g, u, v, w, occ, dt, start, out);
}

int oracle_pressure_anchor(const SmokeSimulationConfig* cfg, const uint8_t* occ) { return int(pressure_anchor(make_geo(cfg), occ)); }

void oracle_pressure_apply(const SmokeSimulationConfig* cfg, const uint8_t* occ, const double* x, double* y) {
    const Geo g = make_geo(cfg);
    const PoissonRows rows = build_rows(g, occ);
    apply_rows(g, rows, x, y);
}

void oracle_pressure_dense(const SmokeSimulationConfig* cfg, const uint8_t* occ, double* a) {
    const Geo g = make_geo(cfg);
    const PoissonRows rows = build_rows(g, occ);
    const size_t n = size_t(g.cells);
    std::fill(a, a + n * n, 0.0);
    for (size_t i = 0; i < n; ++i) {
        a[i * n + i] = rows.diag[i];
        for (int q = 0; q < 6 && rows.cols[i * 6 + q] >= 0; ++q) a[i * n + size_t(rows.cols[i * 6 + q])] = -1.0;
    }
}

void oracle_pressure_rhs(const SmokeSimulationConfig* cfg, const uint8_t* occ, const double* us, const double* vs, const double* ws, double dt, double* rhs) {
    const Geo g = make_geo(cfg);
    pressure_rhs(g, occ, pressure_anchor(g, occ), us, vs, ws, dt, rhs);
}

void oracle_nearest_fluid(const SmokeSimulationConfig* cfg, const uint8_t* occ, int64_t* index) {
    const Geo g = make_geo(cfg);
    for_each_cell(g, [&](int x, int y, int z) {
        const int64_t i = cidx(g, x, y, z);
        index[i] = occ[i] ? nearest_fluid(g, *g.density_bc, occ, x, y, z) : i;
    });
}

namespace {

// FP32 state for the CPU baseline (host port of the step, FP64 reductions).
struct F32Run {
    Geo g;
    size_t n;
    std::vector<float> u, v, w, us, vs, ws, rho, T, rs, ts, zero, tsol, rhs, fx, fy, fz, cx, cy, cz, ox, oy, oz, om, dv, vm;
    std::vector<double> pressure, d_rs, d_ts;
    std::vector<uint8_t> occ;
    Fields<float> f{};
    explicit F32Run(const SmokeSimulationConfig* cfg) : g(make_geo(cfg)), n(size_t(g.cells)) {
        u.assign(size_t(g.faces[0]), 0.f);
        v.assign(size_t(g.faces[1]), 0.f);
        w.assign(size_t(g.faces[2]), 0.f);
        us = u;
        vs = v;
        ws = w;
        for (auto* a : {&rho, &rs, &ts, &zero, &rhs, &fx, &fy, &fz, &cx, &cy, &cz, &ox, &oy, &oz, &om, &dv, &vm}) a->assign(n, 0.f);
        T.assign(n, cfg->ambient_temperature);
        tsol.assign(n, cfg->ambient_temperature);
        pressure.assign(n, 0.0);
        d_rs.assign(n, 0.0);
        d_ts.assign(n, 0.0);
        occ.assign(n, 0);
        f.u = u.data(); f.v = v.data(); f.w = w.data(); f.us = us.data(); f.vs = vs.data(); f.ws = ws.data();
        f.rho = rho.data(); f.T = T.data(); f.rho_src = rs.data(); f.T_src = ts.data();
        f.fux = zero.data(); f.fuy = zero.data(); f.fuz = zero.data();
        f.svx = zero.data(); f.svy = zero.data(); f.svz = zero.data();
        f.T_solid = tsol.data(); f.rhs = rhs.data();
        f.fx = fx.data(); f.fy = fy.data(); f.fz = fz.data(); f.cx = cx.data(); f.cy = cy.data(); f.cz = cz.data();
        f.ox = ox.data(); f.oy = oy.data(); f.oz = oz.data(); f.om = om.data(); f.div = dv.data(); f.vmag = vm.data();
        f.pressure = pressure.data();
        f.occ = occ.data();
    }
    void sources(const SmokeSimulationConfig* cfg, const OraclePlumeSource* src, uint32_t step) {
        if (!src) return;
        oracle_plume_sources(cfg, src, step, d_rs.data(), d_ts.data());
        for (size_t i = 0; i < n; ++i) {
            rs[i] = float(d_rs[i]);
            ts[i] = float(d_ts[i]);
        }
    }
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

}  // namespace

// Whole steps of the FP32 host port, each timed with steady_clock: `warmup`
// untimed steps, then `steps` timed ones (seconds and pressure iterations per
// timed step), optionally on an occupancy grid (uint8, cells).
int oracle_f32_steps(const SmokeSimulationConfig* cfg, const OracleOptions* opt, const OraclePlumeSource* src, const uint8_t* occupancy,
                     int warmup, int steps, int threads, double* step_seconds, int* step_iterations, OracleStats* stats) {
    set_threads(threads);
    F32Run run(cfg);
    if (occupancy) std::copy(occupancy, occupancy + run.n, run.occ.begin());
    OracleStats local{};
    float dt = cfg->dt;
    for (int step = 0; step < warmup + steps; ++step) {
        run.sources(cfg, src, uint32_t(step));
        const auto t0 = std::chrono::steady_clock::now();
        run_stages(cfg, opt, run.f, 0, 4, &local, dt);
        const double secs = seconds_since(t0);
        if (step >= warmup) {
            if (step_seconds) step_seconds[step - warmup] = secs;
            if (step_iterations) step_iterations[step - warmup] = local.iterations;
        }
        if (std::getenv("SMOKE_ORACLE_PROGRESS")) {
            std::fprintf(stderr, "cpu port step %d%s: %.3f s, %d pressure iterations\n", step, step < warmup ? " (warm-up)" : "", secs,
                         local.iterations);
            std::fflush(stderr);
        }
    }
    if (stats) *stats = local;
    return 0;
}

int oracle_f32_benchmark(const SmokeSimulationConfig* cfg, const OracleOptions* opt, const OraclePlumeSource* src, int warmup, int steps, int threads, double* seconds, OracleStats* stats) {
    set_threads(threads);
    F32Run run(cfg);
    OracleStats local{};
    float dt = cfg->dt;
    double elapsed = 0.0;
    for (int step = 0; step < warmup + steps; ++step) {
        run.sources(cfg, src, uint32_t(step));
        const auto t0 = std::chrono::steady_clock::now();
        run_stages(cfg, opt, run.f, 0, 4, &local, dt);
        if (step >= warmup) elapsed += seconds_since(t0);
    }
    if (seconds) *seconds = elapsed;
    if (stats) *stats = local;
    return 0;
}

// Bounded CPU sample of one step from the initial state, timed by stage:
// everything but the pressure iterations once, and `cg_iterations` plain CG
// iterations for the per-iteration cost (the pressure stage time divided by
// the iteration count).  Stage costs do not depend on the flow state.
int oracle_f32_step_sample(const SmokeSimulationConfig* cfg, const OracleOptions* opt, const OraclePlumeSource* src, int cg_iterations, int threads, double* t_other, double* t_iteration) {
    set_threads(threads);
    F32Run run(cfg);
    OracleStats local{};
    float dt = cfg->dt;
    run.sources(cfg, src, 0);
    SmokeSimulationConfig fixed = *cfg;
    fixed.pressure_iterations = cg_iterations > 0 ? cg_iterations : 1;
    OracleOptions fo = *opt;
    fo.pressure_mode = 0;
    auto t0 = std::chrono::steady_clock::now();
    run_stages(&fixed, &fo, run.f, 0, 1, &local, dt);
    double other = seconds_since(t0);
    t0 = std::chrono::steady_clock::now();
    run_stages(&fixed, &fo, run.f, 2, 2, &local, dt);
    const double solve = seconds_since(t0);
    t0 = std::chrono::steady_clock::now();
    run_stages(&fixed, &fo, run.f, 3, 4, &local, dt);
    other += seconds_since(t0);
    if (t_other) *t_other = other;
    if (t_iteration) *t_iteration = solve / fixed.pressure_iterations;
    return 0;
}

}  // extern "C"
