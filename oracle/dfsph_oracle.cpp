// CPU oracle of the DFSPH liquid step -- TEST INFRASTRUCTURE ONLY.
//
// A serial C++ restatement of the reference's DFSPH step (alloy/solvers/sph, Warp kernels run by Warp's
// single-threaded CPU backend), kernel by kernel, in the reference's launch order and arithmetic order. It is
// used by tests/ (parity checker), by __graft_entry__.smoke() (checker) and by bench.py's cpu_baseline leg (a
// timed baseline); nothing in the product (alloy_amd/) links, loads or calls it.
//
// Instantiated for float (one-step parity with the HIP path: same operation order, FMA contraction off, so
// per-particle fields agree bit for bit when inputs agree) and for double (the FP64 density-error oracle).
//
// Pinning (tests/test_oracle_*.py): kernels by closed forms (integral of W, finite-difference gradient, Psi
// against quadrature and its special values); neighbour lists by brute force; reductions against numpy float64;
// density/alpha by the rest-lattice identities (ratio 1 in the bulk and at flat walls, alpha = 1/lattice
// diagonal); solves by residual <= tolerance, kappa >= lower bound, CG/Jacobi agreement and resting-column
// physics; boundary slots by flat-wall, corner and facing-wall special cases; particle walls by the colour
// inversion round trip; advect/project by x + v dt and the push-out invariant phi >= eps; emitters and ledgers by
// the reference's exact-ledger tests; the implicit viscosity by a shear mode's decay against nu k^2, rigid motion
// left untouched and zero viscosity as the identity; its wall term by Nusselt film drainage and the walls'
// impulse balancing the liquid's momentum; CSF by a lattice drop's curvature 2 / R, Laplace pressure and small
// clusters gaining no energy (tests/test_oracle_viscosity.py, test_oracle_wall_viscosity.py, test_oracle_csf.py).
//
// Differences from the reference that are deliberate and shared with the HIP path:
//  - the hash grid is a per-world dense cell grid (cell side h (1 + 1e-5), clamped at its edges). The neighbour
//    SET is grid independent (neighbors.py:187-192 keeps j with fp32 dot(d, d) < h^2); particle-wall colour sums
//    visit cells in z, y, x order and particles in ascending id (the reference's order is Warp's hash order).
//  - per-world float64 sums use a fixed 256-wide tree per chunk and per world (reduce.py:58-164 does the same
//    on CUDA with wp.tile_sum; its CPU path sums serially); both orders are deterministic and batch invariant.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../include/dfsph_abi.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

constexpr int ACTIVE = 1;
constexpr int GEO_PLANE = 1, GEO_SPHERE = 3, GEO_BOX = 7;
constexpr int CHUNK = 256;

template <class T>
struct V3 {
    T x, y, z;
};

template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) {
    T l = length(a);
    if (l > T(0)) return a / l;
    return {T(0), T(0), T(0)};
}
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T tmax(T a, T b) { return a > b ? a : b; }
template <class T> inline T tmin(T a, T b) { return a < b ? a : b; }
template <class T> inline T tclamp(T v, T lo, T hi) { return tmin(tmax(v, lo), hi); }

template <class T> inline T kpi() { return T(3.14159265358979323846); }

// ------------------------------------------------------------------------------------ kernels (kernels.py:41-138)
template <class T>
T kernel_w(T r, T h, int kind) {
    T q = r / h;
    if (q >= T(1)) return T(0);
    T h3 = h * h * h;
    if (kind == 1) {
        T s = T(1) - q;
        return (T(21) / (T(2) * kpi<T>() * h3)) * s * s * s * s * (T(1) + T(4) * q);
    }
    T k = T(8) / (kpi<T>() * h3);
    if (q <= T(0.5)) return k * (T(6) * q * q * q - T(6) * q * q + T(1));
    T s = T(1) - q;
    return k * T(2) * s * s * s;
}

template <class T>
V3<T> kernel_grad(V3<T> rv, T h, int kind) {
    T r = length(rv);
    T q = r / h;
    if (q >= T(1) || r <= T(1.0e-9) * h) return {T(0), T(0), T(0)};
    T h3 = h * h * h;
    T dwdq;
    if (kind == 1) {
        T s = T(1) - q;
        dwdq = (T(21) / (T(2) * kpi<T>() * h3)) * (T(-20) * q * s * s * s);
    } else {
        T k = T(8) / (kpi<T>() * h3);
        if (q <= T(0.5)) {
            dwdq = k * (T(18) * q * q - T(12) * q);
        } else {
            T s = T(1) - q;
            dwdq = k * (T(-6) * s * s);
        }
    }
    return rv * (dwdq / (r * h));
}

template <class T>
T projected_p(T q, int kind) {
    if (q >= T(1)) return T(0);
    T q2 = q * q;
    if (kind == 1) return T(1.5) + q2 * (T(-10.5) + q2 * (T(52.5) + q * (T(-84) + q * (T(52.5) - T(12) * q))));
    if (q <= T(0.5)) return T(1.4) + q2 * (T(-8) + q2 * (T(24) - T(19.2) * q));
    return T(1.6) + q2 * (T(-16) + q * (T(32) + q * (T(-24) + T(6.4) * q)));
}

template <class T>
T g0(T y, int kind) {
    T y2 = y * y;
    if (kind == 1) return y * (T(1.5) + y2 * (T(-3.5) + y2 * (T(10.5) + y * (T(-14) + y * (T(7.5) - T(1.5) * y)))));
    if (y <= T(0.5)) return y * (T(1.4) + y2 * (T(-8) / T(3) + y2 * (T(4.8) - T(3.2) * y)));
    return T(-1) / T(30) + y * (T(1.6) + y2 * (T(-16) / T(3) + y * (T(8) + y * (T(-4.8) + y * (T(16) / T(15))))));
}

template <class T>
T g1(T y, int kind) {
    T y2 = y * y;
    if (kind == 1)
        return y2 * (T(0.75) + y2 * (T(-2.625) + y2 * (T(8.75) + y * (T(-12) + y * (T(6.5625) - (T(4) / T(3)) * y)))));
    if (y <= T(0.5)) return y2 * (T(0.7) + y2 * (T(-2) + y2 * (T(4) - (T(19.2) / T(7)) * y)));
    return T(-1) / T(280) + y2 * (T(0.8) + y2 * (T(-4) + y * (T(6.4) + y * (T(-4) + y * (T(6.4) / T(7))))));
}

// (Psi, dPsi/dd) of a solid half-space at signed distance d with mean curvature H (kernels.py:122-138)
template <class T>
void halfspace_volume(T d, T h, T H, int kind, T& psi, T& dpsi) {
    if (d >= h) {
        psi = T(0);
        dpsi = T(0);
        return;
    }
    T z = tclamp(-d / h, T(-1), T(1));
    T y = std::fabs(z);
    T sgn = T(1);
    if (z < T(0)) sgn = T(-1);
    T g1_full = g1(T(1), kind);
    psi = T(0.5) + sgn * g0(y, kind) - T(0.5) * H * h * (g1(y, kind) - g1_full);
    dpsi = T(0);
    if (d > -h) dpsi = -projected_p(y, kind) * (T(1) - T(0.5) * H * h * z) / h;
}

// ------------------------------------------------------------------------------------ transforms and SDFs
template <class T>
struct Xform {
    V3<T> p;
    T qx, qy, qz, qw;
};

template <class T>
V3<T> quat_rotate(const Xform<T>& X, V3<T> v) {
    V3<T> u{X.qx, X.qy, X.qz};
    V3<T> c = cross(u, v);
    return v * (T(2) * X.qw * X.qw - T(1)) + c * X.qw * T(2) + u * dot(u, v) * T(2);
}

template <class T>
V3<T> quat_rotate_inv(const Xform<T>& X, V3<T> v) {
    Xform<T> I{X.p, -X.qx, -X.qy, -X.qz, X.qw};
    return quat_rotate(I, v);
}

template <class T> V3<T> xf_point(const Xform<T>& X, V3<T> v) { return quat_rotate(X, v) + X.p; }
template <class T> V3<T> xf_vector(const Xform<T>& X, V3<T> v) { return quat_rotate(X, v); }
// transform_point(transform_inverse(X), v): rotate by q^-1, then add -rotate(q^-1, p)
template <class T> V3<T> xf_inv_point(const Xform<T>& X, V3<T> v) {
    return quat_rotate_inv(X, v) + (-quat_rotate_inv(X, X.p));
}

template <class T> inline T sgn1(T v) { return v < T(0) ? T(-1) : T(1); }

// Newton's analytic SDFs (alloy/_vendor/newton/_src/geometry/kernels.py:277-370, soft_contacts_sdf.py:73-100)
template <class T>
void shape_sdf(int geo, V3<T> s, V3<T> p, T& phi, V3<T>& grad) {
    if (geo == GEO_SPHERE) {
        phi = length(p) - s.x;
        T l = length(p);
        grad = l > T(1.0e-8) ? p / l : v3<T>(0, 0, 1);
        return;
    }
    if (geo == GEO_PLANE) {
        phi = p.z;
        grad = v3<T>(0, 0, 1);
        return;
    }
    T qx = std::fabs(p.x) - s.x, qy = std::fabs(p.y) - s.y, qz = std::fabs(p.z) - s.z;
    V3<T> e{tmax(qx, T(0)), tmax(qy, T(0)), tmax(qz, T(0))};
    phi = length(e) + tmin(tmax(qx, tmax(qy, qz)), T(0));
    if (qx > T(0) || qy > T(0) || qz > T(0)) {
        V3<T> c{tclamp(p.x, -s.x, s.x), tclamp(p.y, -s.y, s.y), tclamp(p.z, -s.z, s.z)};
        grad = normalize(p - c);
        return;
    }
    if ((qx > qy && qx > qz) || (qy == T(0) && qz == T(0))) {
        grad = v3<T>(sgn1(p.x), 0, 0);
        return;
    }
    if ((qy > qx && qy > qz) || (qx == T(0) && qz == T(0))) {
        grad = v3<T>(0, sgn1(p.y), 0);
        return;
    }
    grad = v3<T>(0, 0, sgn1(p.z));
}

// ------------------------------------------------------------------------------------ the oracle solver
template <class T>
struct Oracle {
    // parameters (each cast once from the host doubles, as a float kernel argument would be)
    T h, spacing, r, n_full, tot_min, min_nb, omega, warm_den, warm_div, xsph, v_sep;
    T offset, amplitude, taper0, taper1, reach, eps, sweep_len, lattice_grad;
    double relief_speed, rest_speed, flow_ratio, tol_den, tol_div, tol_min, err_factor, lattice_grad_d;
    int kind, solver, min_it, max_it, cap, div_solve, two_sided, friction, curvature, one_way = 0;
    // Weiler 2018 implicit viscosity (viscosity.py)
    int visc_on = 0, visc_max_it = 0;
    T visc_d = T(0);
    float visc_tol = 0.0f;
    std::vector<T> mu, vb, vx, vr, vz, vp, vap, vminv;
    std::vector<int> v_conv, v_iters;
    std::vector<double> v_rz, v_bb, v_step;
    // its wall term (wall_viscosity.py)
    T wall_coeff = T(0), mu_b = T(0);
    std::vector<T> wall_m, wall_w, vshift, vraw, vgk;
    // continuum surface force (csf.py, csf_sparse.py)
    int csf_on = 0;
    T csf_cs = T(0), csf_inv_i = T(0), csf_lo = T(0), csf_hi = T(0), csf_kcoh = T(0);
    std::vector<T> sigma, gamma_, sh_cos, sh_sin;
    std::vector<T> t_mcount, t_kraw, t_ktmp, t_kappa, t_grad, t_g1, t_g2, t_nhat, t_nwall, t_anorm;
    std::vector<int> t_surf, t_wslot;

    int n = 0, W = 0, S = 0, NW = 0;  // NW: wall particles
    std::vector<int> ws;              // world start [W+1]
    std::vector<T> mass, rho0, gravity, wvol;
    std::vector<int> flags, world;
    std::vector<int8_t> phase;
    // shapes
    std::vector<int> sh_start, sh_ids, sh_geo;
    std::vector<V3<T>> sh_scale;
    std::vector<Xform<T>> sh_x;
    std::vector<T> sh_radius, sh_mu;
    // grid
    T cell;
    std::vector<T> g_org;
    std::vector<int> g_dims, cell_base;  // cell_base[w]: first cell of world w (fluid); walls after all fluid
    int n_cells = 0;
    std::vector<int> cell_start, cell_items;
    // neighbour rows
    int wide = 0;
    std::vector<uint16_t> nbr;
    std::vector<uint8_t> nbr_hi;
    std::vector<int> count, group;
    // slots
    std::vector<int> s_shape;
    std::vector<T> s_psi, s_gpsi, s_xb, s_vb, s_phi, s_xbl, s_nbl, s_imp;
    // per particle
    std::vector<T> ratio, alpha, neff, active, reg, src, kappa, acc, err, lower, kv, kd, vtmp;
    // CG
    std::vector<T> pd, z, wv, ap, corr, pair;
    std::vector<int> act;
    // per world
    std::vector<int> converged, iters, restart;
    std::vector<T> err_mean, err_max, tol_scale, tol_one;
    std::vector<double> n_active, err_sum, rz, stepb, sums, flow_z, flow_xy, head;
    std::vector<int> st_overflow, st_projected, st_swept, st_nonfinite, it_den, it_div;
    std::vector<T> st_max_speed, st_den_err, st_div_err, st_den_err_max, st_den_tol;
    std::vector<int64_t> meter;
    std::vector<int> wcidx;            // compact index of each wall particle (ascending ids), -1 otherwise
    std::vector<int64_t> pw_exch, pw_work;  // [NW*3] exchange window, [W] walls' work (2^50 fixed point)
    // emitters, sinks, escape box (emitters.py:32-255) and the int64 ledgers (runtime.py:93-107)
    int E = 0, has_bounds = 0, emit_active = 0;
    V3<T> lo{}, hi{};
    std::vector<int> e_world, off_start, pool_start, pool_ids, cursor, enabled, sink_start;
    std::vector<Xform<T>> e_xform, s_xform;
    std::vector<V3<T>> e_vel, off_local, s_half;
    std::vector<T> e_interval, t_acc;
    std::vector<int64_t> mass0, mass_in, mass_out, mass_act, flux_e, flux_g;
    std::vector<int> escaped, emit_overflow;
    // chunks
    std::vector<int> ch_begin, ch_end, w_chunk;
    // state (bound per call)
    T* q = nullptr;
    T* qd = nullptr;
    const T* f = nullptr;

    V3<T> qv(int i) const { return {q[3 * i], q[3 * i + 1], q[3 * i + 2]}; }
    V3<T> vv(int i) const { return {qd[3 * i], qd[3 * i + 1], qd[3 * i + 2]}; }
    void setq(int i, V3<T> a) { q[3 * i] = a.x; q[3 * i + 1] = a.y; q[3 * i + 2] = a.z; }
    void setv(int i, V3<T> a) { qd[3 * i] = a.x; qd[3 * i + 1] = a.y; qd[3 * i + 2] = a.z; }
    static V3<T> get3(const std::vector<T>& a, size_t k) { return {a[3 * k], a[3 * k + 1], a[3 * k + 2]}; }
    static void put3(std::vector<T>& a, size_t k, V3<T> v) { a[3 * k] = v.x; a[3 * k + 1] = v.y; a[3 * k + 2] = v.z; }

    bool is_fluid(int i) const { return (flags[i] & ACTIVE) != 0 && int(phase[i]) >= 1; }

    int nbr_j(int i, int k) const {
        int b = ws[world[i]];
        int j = b + int(nbr[size_t(k) * n + i]);
        if (wide) j += int(nbr_hi[size_t(k) * n + i]) << 16;
        return j;
    }

    // ---------------------------------------------------------------------------- setup
    int init(const DfsphParams* P, const DfsphLayout* L, const DfsphModel* M) {
        h = T(P->h);
        spacing = T(P->spacing);
        r = T(P->particle_radius);
        n_full = T(P->n_full);
        tot_min = T(P->tot_min);
        min_nb = T(P->min_neighbors_divergence);
        omega = T(P->jacobi_omega);
        warm_den = T(P->warm_start_density);
        warm_div = T(P->warm_start_divergence);
        xsph = T(P->xsph_c);
        v_sep = T(P->contact_separation_speed);
        offset = T(P->wall_offset);
        amplitude = T(P->wall_amplitude);
        taper0 = T(P->taper0);
        taper1 = T(P->taper1);
        reach = T(P->reach);
        eps = T(P->project_eps);
        sweep_len = T(P->sweep_len);
        one_way = P->one_way != 0;
        visc_on = P->viscosity_implicit != 0;
        visc_d = T(P->visc_d);
        visc_tol = float(P->visc_tol);
        visc_max_it = P->visc_max_iterations;
        wall_coeff = visc_on ? T(P->visc_wall_coeff) : T(0);
        mu_b = T(P->boundary_viscosity);
        csf_on = P->csf != 0;
        csf_cs = T(P->csf_c_surface);
        csf_inv_i = T(P->csf_inv_i);
        csf_lo = T(P->csf_band_lo);
        csf_hi = T(P->csf_band_hi);
        csf_kcoh = T(P->csf_k_cohesion);
        lattice_grad = T(P->lattice_grad);
        lattice_grad_d = P->lattice_grad;
        relief_speed = P->compression_relief_speed;
        rest_speed = P->rest_jitter_speed;
        flow_ratio = P->rest_jitter_flow_ratio;
        tol_den = P->max_density_error;
        tol_div = P->max_divergence_error;
        tol_min = std::min(P->min_density_error, P->max_density_error);
        err_factor = P->max_error_factor;
        kind = P->kernel;
        solver = P->pressure_solver;
        min_it = P->min_iterations;
        max_it = P->max_iterations;
        cap = P->neighbor_capacity;
        div_solve = P->divergence_solve;
        two_sided = P->divergence_two_sided;
        friction = P->wall_friction;
        curvature = P->curvature;
        if (cap < 8 || cap > 64) return fail(DFSPH_EINVAL, "neighbor_capacity must be in [8, 64]");
        n = L->n;
        W = L->world_count;
        S = L->shape_count;
        NW = L->wall_particle_count;
        if (n < 0 || W < 1) return fail(DFSPH_EINVAL, "need n >= 0 and world_count >= 1");
        ws.assign(L->world_start, L->world_start + W + 1);
        int maxw = 0;
        for (int w = 0; w < W; ++w) maxw = std::max(maxw, ws[w + 1] - ws[w]);
        wide = maxw > 65536;
        mass.resize(n);
        rho0.resize(n);
        wvol.resize(n);
        flags.assign(M->flags, M->flags + n);
        phase.assign(M->phase, M->phase + n);
        world.assign(M->world, M->world + n);
        for (int i = 0; i < n; ++i) {
            mass[i] = T(M->mass[i]);
            rho0[i] = T(M->rest_density[i]);
            wvol[i] = M->wall_volume ? T(M->wall_volume[i]) : T(0);
        }
        gravity.resize(3 * W);
        for (int k = 0; k < 3 * W; ++k) gravity[k] = T(M->gravity[k]);
        sh_start.assign(L->shape_start, L->shape_start + W + 1);
        sh_ids.assign(L->shape_ids, L->shape_ids + sh_start[W]);
        sh_geo.resize(S);
        sh_scale.resize(S);
        sh_x.resize(S);
        sh_radius.resize(S);
        sh_mu.resize(S);
        for (int s = 0; s < S; ++s) {
            sh_geo[s] = L->shape_geo[s];
            sh_scale[s] = {T(L->shape_scale[3 * s]), T(L->shape_scale[3 * s + 1]), T(L->shape_scale[3 * s + 2])};
            const float* x = L->shape_xform + 7 * s;
            sh_x[s] = {{T(x[0]), T(x[1]), T(x[2])}, T(x[3]), T(x[4]), T(x[5]), T(x[6])};
            sh_radius[s] = T(L->shape_radius[s]);
            sh_mu[s] = T(L->shape_mu[s]);
        }
        // grid: fluid cells of every world, then wall cells of every world
        cell = T((P->h + std::max(P->skin, 0.0)) * (1.0 + 1.0e-5));  // the product's grid (wider with a skin)
        g_org.resize(3 * W);
        g_dims.resize(3 * W);
        cell_base.resize(2 * W + 1);
        long long tot = 0;
        for (int w = 0; w < W; ++w) {
            for (int a = 0; a < 3; ++a) {
                g_org[3 * w + a] = T(L->grid_origin[3 * w + a]);
                g_dims[3 * w + a] = std::max(1, L->grid_dims[3 * w + a]);
            }
        }
        for (int g = 0; g < 2 * W; ++g) {
            int w = g % W;
            cell_base[g] = int(tot);
            tot += (long long)g_dims[3 * w] * g_dims[3 * w + 1] * g_dims[3 * w + 2];
        }
        if (tot > (1LL << 30)) return fail(DFSPH_ECAPACITY, "cell grid too large");
        n_cells = int(tot);
        cell_base[2 * W] = n_cells;
        // buffers
        size_t N = size_t(std::max(n, 1));
        nbr.assign(size_t(cap) * N, 0);
        nbr_hi.assign(wide ? size_t(cap) * N : 1, 0);
        count.assign(N, 0);
        group.assign(N, -1);
        s_shape.assign(4 * N, -1);
        for (auto* v : {&s_psi, &s_phi}) v->assign(4 * N, T(0));
        for (auto* v : {&s_gpsi, &s_xb, &s_vb, &s_xbl, &s_nbl, &s_imp}) v->assign(12 * N, T(0));
        for (auto* v : {&ratio, &alpha, &neff, &active, &reg, &src, &kappa, &err, &lower, &kv, &kd, &pd, &z, &wv,
                        &ap, &corr})
            v->assign(N, T(0));
        acc.assign(3 * N, T(0));
        vtmp.assign(3 * N, T(0));
        pair.assign(2 * N, T(0));
        act.assign(N, 0);
        for (auto* v : {&converged, &iters, &restart, &st_overflow, &st_projected, &st_swept, &st_nonfinite, &it_den,
                        &it_div})
            v->assign(W, 0);
        for (auto* v : {&err_mean, &err_max, &st_max_speed, &st_den_err, &st_div_err, &st_den_err_max, &st_den_tol})
            v->assign(W, T(0));
        tol_scale.assign(W, T(1));
        tol_one.assign(W, T(1));
        for (auto* v : {&n_active, &err_sum, &rz, &stepb}) v->assign(W, 0.0);
        for (auto* v : {&sums, &flow_z, &flow_xy, &head}) v->assign(2 * W, 0.0);
        meter.assign(3 * size_t(std::max(S, 1)), 0);
        for (auto* v : {&mass0, &mass_in, &mass_out, &mass_act, &flux_e, &flux_g}) v->assign(W, 0);
        sh_cos.assign(std::max(S, 1), T(0));
        sh_sin.assign(std::max(S, 1), T(-1));
        if (L->shape_cos && L->shape_sin)
            for (int s = 0; s < S; ++s) {
                sh_cos[s] = T(L->shape_cos[s]);
                sh_sin[s] = T(L->shape_sin[s]);
            }
        if (csf_on) {
            if (!M->surface_tension || !M->csf_gamma) return fail(DFSPH_EINVAL, "csf needs surface_tension and csf_gamma");
            sigma.assign(std::max(n, 1), T(0));
            gamma_.assign(std::max(n, 1), T(0));
            for (int i = 0; i < n; ++i) {
                sigma[i] = T(M->surface_tension[i]);
                gamma_[i] = T(M->csf_gamma[i]);
            }
            for (auto* v : {&t_mcount, &t_kraw, &t_ktmp, &t_kappa}) v->assign(std::max(n, 1), T(0));
            for (auto* v : {&t_grad, &t_g1, &t_g2, &t_nhat, &t_nwall, &t_anorm}) v->assign(3 * size_t(std::max(n, 1)), T(0));
            t_surf.assign(std::max(n, 1), 0);
            t_wslot.assign(std::max(n, 1), -1);
        }
        mu.assign(std::max(n, 1), T(0));
        if (M->viscosity)
            for (int i = 0; i < n; ++i) mu[i] = T(M->viscosity[i]);
        if (visc_on) {
            for (auto* v : {&vb, &vx, &vr, &vz, &vp, &vap}) v->assign(3 * size_t(std::max(n, 1)), T(0));
            vminv.assign(9 * size_t(std::max(n, 1)), T(0));
            v_conv.assign(W, 0);
            v_iters.assign(W, 0);
            for (auto* v : {&v_rz, &v_bb, &v_step}) v->assign(W, 0.0);
            if (wall_coeff > T(0)) {
                wall_m.assign(9 * size_t(std::max(n, 1)), T(0));
                for (auto* v : {&wall_w, &vshift, &vraw}) v->assign(3 * size_t(std::max(n, 1)), T(0));
                vgk.assign(std::max(n, 1), T(0));
            }
        }
        wcidx.assign(std::max(n, 1), -1);
        int nw = 0;
        for (int i = 0; i < n; ++i)
            if (wvol[i] > T(0)) wcidx[i] = nw++;
        pw_exch.assign(3 * size_t(std::max(nw, 1)), 0);
        pw_work.assign(W, 0);
        escaped.assign(W, 0);
        emit_overflow.assign(W, 0);
        sink_start.assign(W + 1, 0);
        // chunk table (reduce.py:167-182)
        w_chunk.push_back(0);
        for (int w = 0; w < W; ++w) {
            for (int b = ws[w]; b < ws[w + 1]; b += CHUNK) {
                ch_begin.push_back(b);
                ch_end.push_back(std::min(b + CHUNK, ws[w + 1]));
            }
            w_chunk.push_back(int(ch_begin.size()));
        }
        active_mass();  // runtime.py:132-134 init_mass
        mass0 = mass_act;
        return DFSPH_OK;
    }

    // ---------------------------------------------------------------------------- reductions (reduce.py)
    // 256-wide fixed tree: a[t] += a[t + s] for s = 128, 64, ..., 1
    static double tree(double* a) {
        for (int s = CHUNK / 2; s >= 1; s >>= 1)
            for (int t = 0; t < s; ++t) a[t] += a[t + s];
        return a[0];
    }
    // values: ncomp per particle (1 or 2); out[w*ncomp + c] sums; mx[w]: max of the last component floored at 0
    void reduce(const std::vector<T>& vals, int ncomp, std::vector<double>& out, T* mx) {
        size_t nch = ch_begin.size();
        std::vector<double> part(nch * ncomp), pmax(nch);
        double a[CHUNK];
        for (size_t c = 0; c < nch; ++c) {
            double m = -3.4e38;
            for (int comp = 0; comp < ncomp; ++comp) {
                for (int t = 0; t < CHUNK; ++t) {
                    int i = ch_begin[c] + t;
                    a[t] = i < ch_end[c] ? double(vals[size_t(i) * ncomp + comp]) : 0.0;
                    if (comp == ncomp - 1 && i < ch_end[c]) m = std::max(m, a[t]);
                }
                part[c * ncomp + comp] = tree(a);
            }
            pmax[c] = m;
        }
        for (int w = 0; w < W; ++w) {
            double m = -3.4e38;
            for (int comp = 0; comp < ncomp; ++comp) {
                for (int t = 0; t < CHUNK; ++t) {
                    double s = 0.0;
                    for (int c = w_chunk[w] + t; c < w_chunk[w + 1]; c += CHUNK) s += part[size_t(c) * ncomp + comp];
                    a[t] = s;
                }
                out[size_t(w) * ncomp + comp] = tree(a);
            }
            for (int c = w_chunk[w]; c < w_chunk[w + 1]; ++c) m = std::max(m, pmax[c]);
            if (mx) mx[w] = T(std::max(m, 0.0));
        }
    }

    // ---------------------------------------------------------------------------- grid + neighbours
    int cell_coord(T x, T o, int dim) const {
        T u = (x - o) / cell;
        T fl = std::floor(u);
        if (!(fl >= T(0))) return 0;  // negative or NaN
        if (fl >= T(dim - 1)) return dim - 1;
        return int(fl);
    }
    int cell_of(int g, V3<T> x, int& cx, int& cy, int& cz) const {
        int w = g % W;
        const int* d = &g_dims[3 * w];
        cx = cell_coord(x.x, g_org[3 * w], d[0]);
        cy = cell_coord(x.y, g_org[3 * w + 1], d[1]);
        cz = cell_coord(x.z, g_org[3 * w + 2], d[2]);
        return cell_base[g] + (cx * d[1] + cy) * d[2] + cz;
    }

    // neighbors.py:36-51
    void set_groups() {
        for (int i = 0; i < n; ++i) {
            int g = -1;
            if ((flags[i] & ACTIVE) != 0 && int(phase[i]) >= 1) g = world[i];
            else if (NW > 0 && wvol[i] > T(0) && (flags[i] & ACTIVE) != 0 && world[i] >= 0) g = -2 - world[i];
            group[i] = g;
        }
    }

    void build_grid() {
        std::vector<int> cnt(size_t(n_cells) + 1, 0), key(std::max(n, 1), -1);
        int cx, cy, cz;
        for (int i = 0; i < n; ++i) {
            int g = group[i];
            if (g == -1) continue;
            int gi = g >= 0 ? g : W + (-2 - g);
            key[i] = cell_of(gi, qv(i), cx, cy, cz);
            cnt[key[i] + 1]++;
        }
        for (int c = 0; c < n_cells; ++c) cnt[c + 1] += cnt[c];
        cell_start = cnt;
        cell_items.assign(std::max(cnt[n_cells], 1), 0);
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int i = 0; i < n; ++i)  // ascending id within each cell
            if (key[i] >= 0) cell_items[fill[key[i]]++] = i;
    }

    void build_neighbors() {
        set_groups();
        build_grid();
        T h2 = h * h;
#pragma omp parallel
        {
        std::vector<int> cand;
#pragma omp for schedule(static, 256)
        for (int i = 0; i < n; ++i) {
            int g = group[i];
            if (g < 0) {
                count[i] = 0;
                continue;
            }
            V3<T> xi = qv(i);
            int cx, cy, cz;
            cell_of(g, xi, cx, cy, cz);
            const int* d = &g_dims[3 * g];
            cand.clear();
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int x = cx + dx, y = cy + dy, zz = cz + dz;
                        if (x < 0 || y < 0 || zz < 0 || x >= d[0] || y >= d[1] || zz >= d[2]) continue;
                        int c = cell_base[g] + (x * d[1] + y) * d[2] + zz;
                        for (int e = cell_start[c]; e < cell_start[c + 1]; ++e) {
                            int j = cell_items[e];
                            if (j == i) continue;
                            V3<T> dd = qv(j) - xi;
                            if (dot(dd, dd) < h2) cand.push_back(j);
                        }
                    }
            std::sort(cand.begin(), cand.end());
            int total = int(cand.size());
            int m = std::min(total, cap);
            int base = ws[g];
            for (int k = 0; k < m; ++k) {
                int off = cand[k] - base;
                nbr[size_t(k) * n + i] = uint16_t(off & 65535);
                if (wide) nbr_hi[size_t(k) * n + i] = uint8_t(off >> 16);
            }
            if (total > cap) {
#pragma omp atomic
                st_overflow[g] += 1;
            }
            count[i] = m;
        }
        }
    }

    // ---------------------------------------------------------------------------- particle walls
    // particle_walls.py:65-128: colour c, grad c, mean wall velocity of world w's walls at x
    void pw_sample(V3<T> x, int w, T& c, V3<T>& gc, V3<T>& mv) const {
        c = T(0);
        gc = v3<T>(0, 0, 0);
        mv = v3<T>(0, 0, 0);
        if (NW == 0) return;
        int g = W + w, cx, cy, cz;
        cell_of(g, x, cx, cy, cz);
        const int* d = &g_dims[3 * w];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int X = cx + dx, Y = cy + dy, Z = cz + dz;
                    if (X < 0 || Y < 0 || Z < 0 || X >= d[0] || Y >= d[1] || Z >= d[2]) continue;
                    int cc = cell_base[g] + (X * d[1] + Y) * d[2] + Z;
                    for (int e = cell_start[cc]; e < cell_start[cc + 1]; ++e) {
                        int j = cell_items[e];
                        V3<T> dd = x - qv(j);
                        T wj = wvol[j] * kernel_w(length(dd), h, kind);
                        if (wj > T(0)) {
                            c += wj;
                            gc = gc + wvol[j] * kernel_grad(dd, h, kind);
                            mv = mv + wj * vv(j);
                        }
                    }
                }
        if (c > T(0)) mv = mv / c;
    }
    T colour_distance(T c) const {
        T lo = -h, hi = h;
        for (int k = 0; k < 30; ++k) {
            T mid = T(0.5) * (lo + hi);
            T psi, dpsi;
            halfspace_volume(mid, h, T(0), kind, psi, dpsi);
            if (psi > c) lo = mid;
            else hi = mid;
        }
        return T(0.5) * (lo + hi);
    }
    T pw_candidate(V3<T> x, int w) const {
        T c;
        V3<T> gc, mv;
        pw_sample(x, w, c, gc, mv);
        if (c < T(1.0e-4) || length(gc) <= T(0)) return T(1.0e30);
        return colour_distance(tmin(c, T(1)));
    }

    // ---------------------------------------------------------------------------- boundary slots
    // boundary_slots.py:13-38
    int slot_insert(int i, int nn, int s, T phi, int face) {
        int pos = nn, k = nn - 1;
        while (k >= 0) {
            if (s_phi[4 * i + k] <= phi) break;
            pos = k;
            k -= 1;
        }
        if (pos >= 4) return -1;
        int last = nn >= 4 ? 3 : nn;
        for (k = last; k > pos; --k) {
            s_shape[4 * i + k] = s_shape[4 * i + k - 1];
            s_phi[4 * i + k] = s_phi[4 * i + k - 1];
        }
        s_shape[4 * i + pos] = s;
        s_phi[4 * i + pos] = phi;
        (void)face;
        return pos;
    }
    void slot_copy(int i, int src_, int dst) {
        s_shape[4 * i + dst] = s_shape[4 * i + src_];
        s_phi[4 * i + dst] = s_phi[4 * i + src_];
        put3(s_nbl, 4 * i + dst, get3(s_nbl, 4 * i + src_));
    }
    // boundary_slots.py:50-102
    int shape_insert(int i, int nn, int s, T phi, V3<T> nrm, T near_) {
        if (nn < 4) {
            int rr = slot_insert(i, nn, s, phi, -1);
            if (rr >= 0) {
                for (int k = nn; k > rr; --k) put3(s_nbl, 4 * i + k, get3(s_nbl, 4 * i + k - 1));
                put3(s_nbl, 4 * i + rr, nrm);
                return 1;
            }
            return 0;
        }
        const T cos_same = T(0.85);
        int ev = -1;
        T ev_phi = T(-1);
        for (int m = 0; m < 4; ++m)
            if (phi >= near_ && s_phi[4 * i + m] <= phi && dot(get3(s_nbl, 4 * i + m), nrm) > cos_same) {
                ev = 4;
                ev_phi = phi;
            }
        for (int k = 0; k < 4; ++k) {
            T pk = s_phi[4 * i + k];
            int dup = 0;
            for (int m = 0; m < 4; ++m)
                if (m != k && s_phi[4 * i + m] <= pk && dot(get3(s_nbl, 4 * i + m), get3(s_nbl, 4 * i + k)) > cos_same)
                    dup = 1;
            if (phi <= pk && dot(nrm, get3(s_nbl, 4 * i + k)) > cos_same) dup = 1;
            if (dup == 1 && pk >= near_ && pk > ev_phi) {
                ev = k;
                ev_phi = pk;
            }
        }
        if (ev < 0 && phi < s_phi[4 * i + 3]) ev = 3;
        if (ev < 0 || ev == 4) return 0;
        for (int k = ev; k < 3; ++k) slot_copy(i, k + 1, k);
        int pos = 3;
        while (pos > 0 && s_phi[4 * i + pos - 1] > phi) {
            slot_copy(i, pos - 1, pos);
            pos -= 1;
        }
        s_shape[4 * i + pos] = s;
        s_phi[4 * i + pos] = phi;
        put3(s_nbl, 4 * i + pos, nrm);
        return 1;
    }

    // boundary.py:190-238
    void find_slots(int i) {
        for (int k = 0; k < 4; ++k) s_shape[4 * i + k] = -1;
        if (!is_fluid(i)) return;
        int w = world[i];
        V3<T> x = qv(i);
        int nn = 0;
        for (int c = sh_start[w]; c < sh_start[w + 1]; ++c) {
            int s = sh_ids[c];
            const Xform<T>& X = sh_x[s];
            if (sh_geo[s] != GEO_PLANE)
                if (length(x - X.p) > sh_radius[s] + h) continue;
            V3<T> xl = xf_inv_point(X, x);
            T phi;
            V3<T> gl;
            shape_sdf(sh_geo[s], sh_scale[s], xl, phi, gl);
            if (phi < reach && phi + offset < h) {
                V3<T> nw = normalize(xf_vector(X, gl));
                if (shape_insert(i, nn, s, phi, nw, T(0.1) * h) > 0 && nn < 4) nn += 1;
            }
        }
        if (NW > 0) {
            T phi = pw_candidate(x, w);
            if (phi < reach && phi + offset < h)
                if (slot_insert(i, nn, S, phi, -1) >= 0 && nn < 4) nn += 1;
        }
    }

    void taper(T& v0, T& v1, T phi) const {
        if (phi <= taper0) return;
        T L = taper1 - taper0;
        T t = tmin((phi - taper0) / L, T(1));
        T sm = T(1) - t * t * (T(3) - T(2) * t);
        T dsm = T(6) * t * (t - T(1)) / L;
        T a = v0 * sm, b = v1 * sm + v0 * dsm;
        v0 = a;
        v1 = b;
    }

    T mean_curvature(int geo, V3<T> sc) const {
        if (geo == GEO_SPHERE) {
            T kmax = T(2) / h;
            return T(-2) * tclamp(T(1) / sc.x, -kmax, kmax);
        }
        return T(0);
    }

    // boundary.py:290-333
    void fill_one(int i, int k, V3<T> x, int w) {
        size_t sk = 4 * size_t(i) + k;
        int s = s_shape[sk];
        if (s < 0) {
            s_psi[sk] = T(0);
            put3(s_gpsi, sk, v3<T>(0, 0, 0));
            return;
        }
        if (s == S && NW > 0) {  // particle wall (boundary.py:272-287)
            T c;
            V3<T> gc, mv;
            pw_sample(x, w, c, gc, mv);
            T lg = length(gc);
            V3<T> nrm = v3<T>(0, 0, 1);
            if (lg > T(0)) nrm = (-gc) / lg;
            T phi = colour_distance(tclamp(c, T(0), T(1)));
            V3<T> xb = x - phi * nrm;
            T v0, v1;
            halfspace_volume(phi + offset, h, T(0), kind, v0, v1);
            v0 = v0 * amplitude;
            v1 = v1 * amplitude;
            taper(v0, v1, phi);
            s_phi[sk] = phi;
            s_psi[sk] = v0;
            put3(s_gpsi, sk, v1 * nrm);
            put3(s_xb, sk, xb);
            put3(s_vb, sk, mv);
            put3(s_xbl, sk, xb);
            put3(s_nbl, sk, nrm);
            return;
        }
        const Xform<T>& X = sh_x[s];
        V3<T> xl = xf_inv_point(X, x);
        T phi;
        V3<T> gl;
        shape_sdf(sh_geo[s], sh_scale[s], xl, phi, gl);
        V3<T> nrm = normalize(xf_vector(X, gl));
        V3<T> nl = normalize(gl);
        put3(s_nbl, sk, nl);
        put3(s_xbl, sk, xl - phi * nl);
        T H = T(0);
        if (curvature > 0) H = mean_curvature(sh_geo[s], sh_scale[s]);
        T v0, v1;
        halfspace_volume(phi + offset, h, H, kind, v0, v1);
        v0 = v0 * amplitude;
        v1 = v1 * amplitude;
        taper(v0, v1, phi);
        V3<T> xb = x - phi * nrm;
        s_phi[sk] = phi;
        s_psi[sk] = v0;
        put3(s_gpsi, sk, v1 * nrm);
        put3(s_xb, sk, xb);
        put3(s_vb, sk, v3<T>(0, 0, 0));  // static shapes: point_velocity is zero
    }

    // boundary.py:336-396
    void fill_slots(int i) {
        V3<T> x = qv(i);
        for (int k = 0; k < 4; ++k) fill_one(i, k, x, world[i]);
        T ps[4];
        for (int k = 0; k < 4; ++k) ps[k] = s_psi[4 * i + k];
        for (int b = 0; b < 4; ++b) {
            if (s_shape[4 * i + b] < 0) continue;
            V3<T> gb = get3(s_gpsi, 4 * i + b);
            T ob = T(0), sb = T(0);
            for (int c = 0; c < 4; ++c) {
                if (c == b || s_shape[4 * i + c] < 0) continue;
                V3<T> gc = get3(s_gpsi, 4 * i + c);
                T lb = length(gb), lc = length(gc);
                T cbc = T(0);
                if (lb > T(0) && lc > T(0)) cbc = tclamp(dot(gb, gc) / (lb * lc), T(-1), T(1));
                ob += (T(1) - std::fabs(cbc)) * ps[c];
                if (cbc > T(0) && (ps[c] > ps[b] || (ps[c] == ps[b] && c < b))) sb += cbc;
            }
            ob = tmin(ob, T(1));
            s_psi[4 * i + b] = ps[b] * tmax(T(1) - T(0.5) * ob - sb, T(0));
            put3(s_gpsi, 4 * i + b, gb * tmax(T(1) - ob - sb, T(0)));
        }
    }

    void boundary_slots() {
        // find runs over every particle first (slot_nrm is scratch shared with nbl), then fill
        #pragma omp parallel for schedule(static, 256)
        for (int i = 0; i < n; ++i) find_slots(i);
        #pragma omp parallel for schedule(static, 256)
        for (int i = 0; i < n; ++i) fill_slots(i);
    }

    // ---------------------------------------------------------------------------- density / alpha
    // pressure.py:144-193
    void density_alpha() {
#pragma omp parallel for schedule(static, 256)
        for (int i = 0; i < n; ++i) {
            if (!is_fluid(i)) {
                ratio[i] = alpha[i] = neff[i] = active[i] = reg[i] = T(0);
                continue;
            }
            V3<T> xi = qv(i);
            T mi = mass[i];
            T inv_r0 = T(1) / rho0[i];
            T rho = mi * kernel_w(T(0), h, kind);
            V3<T> G{T(0), T(0), T(0)};
            T sq = T(0);
            int cnt = count[i];
            for (int k = 0; k < cnt; ++k) {
                int j = nbr_j(i, k);
                V3<T> rv = xi - qv(j);
                T mj = mass[j];
                rho += mj * kernel_w(length(rv), h, kind);
                V3<T> g = kernel_grad(rv, h, kind);
                G = G + (mj * inv_r0) * g;
                sq += (mi * mj * inv_r0 * inv_r0) * dot(g, g);
            }
            T psi = T(0), lam = T(0);
            for (int b = 0; b < 4; ++b) {
                if (s_shape[4 * i + b] < 0) continue;
                psi += s_psi[4 * i + b];
                G = G + get3(s_gpsi, 4 * i + b);
                for (int c = b + 1; c < 4; ++c)
                    if (s_shape[4 * i + c] >= 0)
                        lam += tmax(T(-2) * dot(get3(s_gpsi, 4 * i + b), get3(s_gpsi, 4 * i + c)), T(0));
            }
            T tot = dot(G, G) + sq + lam;
            lam += tmax(tot_min - tot, T(0));
            tot = tmax(tot, tot_min);
            reg[i] = lam;
            ratio[i] = rho * inv_r0 + psi;
            alpha[i] = tot > T(1.0e-12) ? T(1) / tot : T(0);
            neff[i] = T(cnt) + psi * n_full;
            active[i] = T(1);
        }
        std::vector<double> out(W);
        reduce(active, 1, out, nullptr);
        n_active = out;
    }

    // ---------------------------------------------------------------------------- solves
    T density_rate(int i) const {
        V3<T> xi = qv(i), vi = vv(i);
        T inv_r0 = T(1) / rho0[i];
        T rate = T(0);
        for (int k = 0; k < count[i]; ++k) {
            int j = nbr_j(i, k);
            V3<T> g = kernel_grad(xi - qv(j), h, kind);
            rate += (mass[j] * inv_r0) * dot(vi - vv(j), g);
        }
        for (int b = 0; b < 4; ++b)
            if (s_shape[4 * i + b] >= 0) rate += dot(vi - get3(s_vb, 4 * i + b), get3(s_gpsi, 4 * i + b));
        return rate;
    }

    // pressure.py:214-261
    void solve_init(int mode, T dt, T warm, T relief, const std::vector<T>& store) {
#pragma omp parallel for schedule(static, 256)
        for (int i = 0; i < n; ++i) {
            if (!is_fluid(i)) {
                src[i] = kappa[i] = lower[i] = T(0);
                continue;
            }
            T rate = density_rate(i);
            if (mode == 0) {
                T kdv = T(0);
                if (two_sided) kdv = tmax(warm_den * tmin(kd[i], T(2.5e-4)) / (dt * dt), T(0));
                T lb = -kdv;
                T rr = rate;
                if (kdv <= T(0)) rr = tmax(rate, T(0));
                if (neff[i] < min_nb) {
                    rr = T(0);
                    lb = T(0);
                }
                src[i] = rr;
                lower[i] = lb;
                kappa[i] = tmax(warm * tmin(store[i], T(0.5)) / dt, lb);
            } else {
                T rs = ratio[i] + dt * rate;
                if (ratio[i] - T(1) > relief) rs = T(1) + relief + dt * rate;
                src[i] = rs;
                lower[i] = T(0);
                kappa[i] = warm * tmin(store[i], T(2.5e-4)) / (dt * dt);
            }
        }
    }

    // pressure.py:264-277
    V3<T> pressure_accel(const std::vector<T>& kap, int i) const {
        V3<T> xi = qv(i);
        T ki = kap[i] / rho0[i];
        V3<T> a{T(0), T(0), T(0)};
        for (int k = 0; k < count[i]; ++k) {
            int j = nbr_j(i, k);
            V3<T> g = kernel_grad(xi - qv(j), h, kind);
            a = a - (mass[j] * (ki + kap[j] / rho0[j])) * g;
        }
        for (int b = 0; b < 4; ++b)
            if (s_shape[4 * i + b] >= 0) a = a - kap[i] * get3(s_gpsi, 4 * i + b);
        return a;
    }

    // pressure.py:294-310
    T accel_divergence(int i) const {
        V3<T> xi = qv(i), ai = get3(acc, i);
        T inv_r0 = T(1) / rho0[i];
        T ar = T(0);
        for (int k = 0; k < count[i]; ++k) {
            int j = nbr_j(i, k);
            V3<T> g = kernel_grad(xi - qv(j), h, kind);
            ar += (mass[j] * inv_r0) * dot(ai - get3(acc, j), g);
        }
        for (int b = 0; b < 4; ++b)
            if (s_shape[4 * i + b] >= 0) ar += dot(ai, get3(s_gpsi, 4 * i + b));
        return ar;
    }

    bool world_skip(int i, int gated) const {
        int w = world[i];
        if (converged[w] != 0) return true;
        return gated != 0 && restart[w] == 0;
    }

    // krylov.py:77-109 (cg_accel + cg_apply): out (+)= A vec (- b)
    void product(int mode, T dt, const std::vector<T>& vec, std::vector<T>& out, int gated, int bias, int accum) {
#pragma omp parallel for schedule(static, 256)
        for (int i = 0; i < n; ++i) {
            if (!is_fluid(i)) {
                put3(acc, i, v3<T>(0, 0, 0));
                continue;
            }
            if (world_skip(i, gated)) continue;
            put3(acc, i, pressure_accel(vec, i));
        }
        T sc = mode == 0 ? dt : dt * dt;
#pragma omp parallel for schedule(static, 256)
        for (int i = 0; i < n; ++i) {
            if (!is_fluid(i)) {
                out[i] = T(0);
                pair[2 * i] = pair[2 * i + 1] = T(0);
                continue;
            }
            if (world_skip(i, gated)) continue;
            T v = -sc * (accel_divergence(i) - reg[i] * vec[i]);
            pair[2 * i] = vec[i] * v;
            pair[2 * i + 1] = T(0);
            if (bias) v -= (mode == 0 ? src[i] : src[i] - T(1));
            if (accum) v += out[i];
            out[i] = v;
        }
    }

    bool stop_test(int w, int it, double mean, T tol, T tol_max, const std::vector<T>& scale, int mi, int ma) const {
        T ts = scale[w];
        return (it >= mi && T(mean) <= tol * ts && err_max[w] <= tol_max * ts) || it >= ma;
    }

    int cg_solve(int mode, T dt, T tol, T tol_max, const std::vector<T>& scale, int mi, int ma) {
        T sc = mode == 0 ? dt : dt * dt;
        std::fill(restart.begin(), restart.end(), 0);
        product(mode, dt, kappa, wv, 0, 1, 0);
#pragma omp parallel for schedule(static, 256)
        for (int i = 0; i < n; ++i) {  // cg_init_kernel
            if (!is_fluid(i)) {
                act[i] = 0;
                z[i] = pd[i] = pair[2 * i] = pair[2 * i + 1] = T(0);
                continue;
            }
            int a = 0;
            T zi = T(0);
            if (kappa[i] > lower[i] || wv[i] < T(0)) {
                a = 1;
                zi = -wv[i] * alpha[i] / sc;
            }
            act[i] = a;
            z[i] = zi;
            pd[i] = zi;
            pair[2 * i] = -wv[i] * zi;
            pair[2 * i + 1] = T(0);
        }
        reduce(pair, 2, sums, nullptr);
        for (int w = 0; w < W; ++w) {
            rz[w] = sums[2 * w];
            restart[w] = 0;
        }
        int n_unconv = 0;
        for (int w = 0; w < W; ++w) n_unconv += converged[w] == 0;
        int loops = 0;
        while (n_unconv != 0) {
            ++loops;
            product(mode, dt, pd, ap, 0, 0, 0);
            reduce(pair, 2, sums, nullptr);
#pragma omp parallel for schedule(static, 256)
            for (int i = 0; i < n; ++i) {  // cg_step_kernel
                if (!is_fluid(i)) {
                    corr[i] = T(0);
                    continue;
                }
                int w = world[i];
                if (converged[w] != 0) continue;
                double pap = sums[2 * w];
                double a64 = 0.0;
                if (pap > 1.0e-30) a64 = rz[w] / pap;
                T a = T(a64);
                T c = T(0);
                if (act[i] != 0) {
                    T k = kappa[i] + a * pd[i];
                    if (k < lower[i]) {
                        c = lower[i] - k;
                        k = lower[i];
                        act[i] = 0;
#pragma omp atomic
                        restart[w] += 1;
                    }
                    kappa[i] = k;
                }
                corr[i] = c;
                wv[i] = wv[i] + a * ap[i];
            }
            product(mode, dt, corr, wv, 1, 0, 1);
#pragma omp parallel for schedule(static, 256)
            for (int i = 0; i < n; ++i) {  // cg_finish_kernel
                if (!is_fluid(i)) {
                    pair[2 * i] = pair[2 * i + 1] = T(0);
                    err[i] = T(0);
                    continue;
                }
                int w = world[i];
                if (converged[w] != 0) continue;
                T wvv = wv[i];
                if (act[i] == 0 && wvv < T(0)) {
                    act[i] = 1;
#pragma omp atomic
                    restart[w] += 1;
                }
                T zi = T(0);
                if (act[i] != 0) zi = -wvv * alpha[i] / sc;
                z[i] = zi;
                T e = tmax(-wvv, T(0));
                if (kappa[i] > lower[i]) e = std::fabs(wvv);
                if (mode == 0) {
                    if (neff[i] < min_nb) e = T(0);
                    e *= dt;
                }
                err[i] = e;
                pair[2 * i] = -wvv * zi;
                pair[2 * i + 1] = e;
            }
            reduce(pair, 2, sums, err_max.data());
            n_unconv = 0;
            for (int w = 0; w < W; ++w) {  // cg_beta_kernel
                if (converged[w] != 0) continue;
                int it = iters[w] + 1;
                iters[w] = it;
                double rz_new = sums[2 * w];
                double mean = sums[2 * w + 1] / std::max(n_active[w], 1.0);
                err_mean[w] = T(mean);
                double beta = 0.0;
                if (restart[w] == 0 && rz[w] > 1.0e-30) beta = rz_new / rz[w];
                restart[w] = 0;
                rz[w] = rz_new;
                stepb[w] = beta;
                if (stop_test(w, it, double(T(mean)), tol, tol_max, scale, mi, ma)) converged[w] = 1;
                else n_unconv += 1;
            }
#pragma omp parallel for schedule(static, 256)
            for (int i = 0; i < n; ++i) {  // cg_direction_kernel
                if (!is_fluid(i)) continue;
                int w = world[i];
                if (converged[w] != 0) continue;
                if (act[i] != 0) pd[i] = z[i] + T(stepb[w]) * pd[i];
                else pd[i] = T(0);
            }
        }
        return loops;
    }

    int jacobi_solve(int mode, T dt, T tol, T tol_max, const std::vector<T>& scale, int mi, int ma) {
        int n_unconv = 0;
        for (int w = 0; w < W; ++w) n_unconv += converged[w] == 0;
        int loops = 0;
        while (n_unconv != 0) {
            ++loops;
            for (int i = 0; i < n; ++i) {  // accel_kernel
                if (!is_fluid(i)) {
                    put3(acc, i, v3<T>(0, 0, 0));
                    continue;
                }
                if (converged[world[i]] != 0) continue;
                put3(acc, i, pressure_accel(kappa, i));
            }
            for (int i = 0; i < n; ++i) {  // jacobi_kernel
                if (!is_fluid(i)) {
                    err[i] = T(0);
                    continue;
                }
                if (converged[world[i]] != 0) continue;
                T ar = accel_divergence(i) - reg[i] * kappa[i];
                T s, scale_;
                if (mode == 0) {
                    ar *= dt;
                    s = -src[i];
                    scale_ = T(1) / dt;
                } else {
                    ar *= dt * dt;
                    s = T(1) - src[i];
                    scale_ = T(1) / (dt * dt);
                }
                T rr = s - ar;
                T lb = lower[i];
                T k_new = tmax(kappa[i] - omega * rr * alpha[i] * scale_, lb);
                kappa[i] = k_new;
                T e = tmax(-rr, T(0));
                if (k_new > lb) e = std::fabs(rr);
                if (mode == 0) {
                    if (neff[i] < min_nb) e = T(0);
                    e *= dt;
                }
                err[i] = e;
            }
            std::vector<double> out(W);
            reduce(err, 1, out, err_max.data());
            err_sum = out;
            n_unconv = 0;
            for (int w = 0; w < W; ++w) {  // check_kernel
                if (converged[w] != 0) continue;
                int it = iters[w] + 1;
                iters[w] = it;
                double mean = err_sum[w] / std::max(n_active[w], 1.0);
                err_mean[w] = T(mean);
                if (stop_test(w, it, double(T(mean)), tol, tol_max, scale, mi, ma)) converged[w] = 1;
                else n_unconv += 1;
            }
        }
        return loops;
    }

    // solve_driver.py:34-82 (resting-jitter rule)
    void flow_speed() {
        if (!(rest_speed > 0.0) || !(flow_ratio > 0.0)) return;
        std::vector<T> ta(2 * size_t(std::max(n, 1))), tb(2 * size_t(std::max(n, 1)));
        for (int i = 0; i < n; ++i) {
            T a0 = 0, a1 = 0, b0 = 0, b1 = 0;
            if (is_fluid(i)) {
                V3<T> v = vv(i);
                a0 = v.z;
                a1 = T(1);
                b0 = v.x;
                b1 = v.y;
            }
            ta[2 * i] = a0;
            ta[2 * i + 1] = a1;
            tb[2 * i] = b0;
            tb[2 * i + 1] = b1;
        }
        reduce(ta, 2, flow_z, nullptr);
        reduce(tb, 2, flow_xy, nullptr);
    }
    void density_tolerance(T dt) {
        if (!(rest_speed > 0.0)) return;
        std::vector<T> terms(2 * size_t(std::max(n, 1)));
        for (int i = 0; i < n; ++i) {
            T k = T(0);
            if (is_fluid(i)) k = tmax(kd[i], T(0));
            terms[2 * i] = k;
            terms[2 * i + 1] = k * k;
        }
        reduce(terms, 2, head, nullptr);
        T speed = T(rest_speed), rat = T(flow_ratio), tol = T(tol_den), tmin_ = T(tol_min);
        for (int w = 0; w < W; ++w) {
            T s = T(1);
            T g = length(v3<T>(gravity[3 * w], gravity[3 * w + 1], gravity[3 * w + 2]));
            T s1 = T(head[2 * w]), s2 = T(head[2 * w + 1]);
            T v = speed;
            double nn = flow_z[2 * w + 1];
            if (rat > T(0) && nn > 0.0) {
                V3<T> u{T(flow_xy[2 * w] / nn), T(flow_xy[2 * w + 1] / nn), T(flow_z[2 * w] / nn)};
                v = tmax(v, rat * length(u));
            }
            if (g > T(0) && s1 > T(0) && s2 > T(0)) {
                T depth = T(1.5) * s2 / (s1 * dt * dt * g);
                s = tclamp(v * dt / (depth * tol), tmin_ / tol, T(1));
            }
            tol_scale[w] = s;
        }
    }

    // solve_driver.py:121-162 plus contact limit, apply pressure, lone friction (pressure.py:393-480)
    void solve(int mode, double dt_host, int fixed_iters, const T* warm_override = nullptr) {
        bool div = mode == 0;
        T dt = T(dt_host);
        std::vector<T>& store = div ? kv : kd;
        double tolh = div ? tol_div : tol_den;
        if (!div) density_tolerance(dt);
        const std::vector<T>& scale = div ? tol_one : tol_scale;
        T warm = div ? warm_div : warm_den;
        T relief = T(1.0e30);
        if (!div && relief_speed > 0.0) relief = T(relief_speed * dt_host * double(lattice_grad_host()));
        solve_init(mode, dt, warm, relief, store);
        if (warm_override)
            for (int i = 0; i < n; ++i)
                if (is_fluid(i)) kappa[i] = tmax(warm_override[i], lower[i]);
        for (int w = 0; w < W; ++w) {  // begin_solve_kernel
            iters[w] = 0;
            err_mean[w] = T(0);
            converged[w] = n_active[w] > 0.0 ? 0 : 1;
        }
        int mi = min_it, ma = max_it;
        if (fixed_iters >= 0) mi = ma = std::max(fixed_iters, 1);
        T tol = T(tolh), tol_max = T(tolh * err_factor);
        if (solver == 0) cg_solve(mode, dt, tol, tol_max, scale, mi, ma);
        else jacobi_solve(mode, dt, tol, tol_max, scale, mi, ma);
        contact_limit(dt);
        apply_pressure(mode, dt, store);
        if (friction) lone_friction(dt);
        for (int w = 0; w < W; ++w) {
            (div ? it_div : it_den)[w] = iters[w];
            (div ? st_div_err : st_den_err)[w] = err_mean[w];
        }
        if (!div) {
            for (int w = 0; w < W; ++w) st_den_err_max[w] = T(0);
            for (int i = 0; i < n; ++i)
                if (is_fluid(i)) st_den_err_max[world[i]] = tmax(st_den_err_max[world[i]], err[i]);
            for (int w = 0; w < W; ++w) st_den_tol[w] = T(tolh) * tol_scale[w];
        }
    }
    double lattice_grad_host() const { return lattice_grad_d; }

    void contact_limit(T dt) {
#pragma omp parallel for schedule(static, 256)
        for (int i = 0; i < n; ++i) {
            if (!is_fluid(i) || count[i] > 0 || kappa[i] <= T(0)) continue;
            V3<T> G{0, 0, 0}, vb{0, 0, 0};
            T wsum = T(0);
            for (int b = 0; b < 4; ++b)
                if (s_shape[4 * i + b] >= 0) {
                    G = G + get3(s_gpsi, 4 * i + b);
                    vb = vb + s_psi[4 * i + b] * get3(s_vb, 4 * i + b);
                    wsum += s_psi[4 * i + b];
                }
            T g = length(G);
            if (g <= T(0)) continue;
            if (wsum > T(0)) vb = vb / wsum;
            T u = dot(vv(i) - vb, (-G) / g);
            kappa[i] = tmin(kappa[i], tmax(v_sep - u, T(0)) / (dt * g));
        }
    }

    void apply_pressure(int mode, T dt, std::vector<T>& store) {
#pragma omp parallel for schedule(static, 256)
        for (int i = 0; i < n; ++i) {
            if (!is_fluid(i)) continue;
            V3<T> a = pressure_accel(kappa, i);
            setv(i, vv(i) + dt * a);
            T mk = mass[i] * kappa[i] * dt;
            for (int b = 0; b < 4; ++b)
                if (s_shape[4 * i + b] >= 0)
                    put3(s_imp, 4 * i + b, get3(s_imp, 4 * i + b) + mk * get3(s_gpsi, 4 * i + b));
            store[i] = mode == 0 ? kappa[i] * dt : kappa[i] * dt * dt;
        }
    }

    void lone_friction(T dt) {
#pragma omp parallel for schedule(static, 256)
        for (int i = 0; i < n; ++i) {
            if (!is_fluid(i) || count[i] > 0 || kappa[i] <= T(0)) continue;
            V3<T> G{0, 0, 0}, vb{0, 0, 0};
            T wsum = T(0), pbest = T(0);
            int best = -1;
            for (int b = 0; b < 4; ++b) {
                int s = s_shape[4 * i + b];
                if (s >= 0) {
                    G = G + get3(s_gpsi, 4 * i + b);
                    vb = vb + s_psi[4 * i + b] * get3(s_vb, 4 * i + b);
                    wsum += s_psi[4 * i + b];
                    if (s_psi[4 * i + b] > pbest) {
                        pbest = s_psi[4 * i + b];
                        best = b;
                    }
                }
            }
            T g = length(G);
            if (g <= T(0) || wsum <= T(0) || best < 0) continue;
            int s = s_shape[4 * i + best];
            if (s >= S) continue;
            T mu = sh_mu[s];
            if (mu <= T(0)) continue;
            V3<T> nn = (-G) / g;
            V3<T> u = vv(i) - vb / wsum;
            V3<T> vt = u - dot(u, nn) * nn;
            T lt = length(vt);
            if (lt <= T(0)) continue;
            V3<T> dv = tmin(T(1), mu * dt * kappa[i] * g / lt) * vt;
            setv(i, vv(i) - dv);
            put3(s_imp, 4 * i + best, get3(s_imp, 4 * i + best) + mass[i] * dv);
        }
    }

    // ---------------------------------------------------------------------------- non-pressure (integrate.py:150-178)
    void nonpressure(T dt) {
#pragma omp parallel for schedule(static, 256)
        for (int i = 0; i < n; ++i) {
            V3<T> vi = vv(i);
            if (!is_fluid(i)) {
                put3(vtmp, i, vi);
                continue;
            }
            int w = world[i];
            V3<T> a{gravity[3 * w], gravity[3 * w + 1], gravity[3 * w + 2]};
            if (f) a = a + V3<T>{f[3 * i], f[3 * i + 1], f[3 * i + 2]} / mass[i];
            V3<T> v = vi + dt * a;
            if (xsph > T(0)) {
                V3<T> xi = qv(i);
                V3<T> cr{0, 0, 0};
                T rho_i = ratio[i] * rho0[i];
                for (int k = 0; k < count[i]; ++k) {
                    int j = nbr_j(i, k);
                    T rho_bar = T(0.5) * (rho_i + ratio[j] * rho0[j]);
                    if (rho_bar > T(0)) {
                        T wk = kernel_w(length(xi - qv(j)), h, kind);
                        cr = cr + (mass[j] / rho_bar * wk) * (vv(j) - vi);
                    }
                }
                v = v + xsph * cr;
            }
            put3(vtmp, i, v);
        }
        for (int i = 0; i < n; ++i) setv(i, get3(vtmp, i));
        if (csf_on) csf_stage(dt);  // solver.py:304-305
    }

    // ---------------------------------------------------------------------------- surface tension (csf.py)
    static T csf_weight(T count, T lo, T hi) {  // csf_sparse.py:44-48
        T x = tclamp((count - lo) / (hi - lo), T(0), T(1));
        return x * x * (T(3) - T(2) * x);
    }
    T cohesion_c(T r) const {  // tension.py:81-92
        if (r > h || r <= T(0)) return T(0);
        T s = (h - r) * (h - r) * (h - r) * r * r * r;
        if (r > T(0.5) * h) return csf_kcoh * s;
        T h3 = h * h * h;
        return csf_kcoh * (T(2) * s - h3 * h3 / T(64));
    }
    void shepard_vec(const std::vector<T>& vin, std::vector<T>& vout) {  // csf.py:83-101
        for (int i = 0; i < n; ++i) {
            V3<T> vi = get3(vin, i);
            if (!is_fluid(i)) {
                put3(vout, i, vi);
                continue;
            }
            V3<T> xi = qv(i);
            T w0 = kernel_w(T(0), h, kind), den = w0;
            V3<T> s = w0 * vi;
            for (int k = 0; k < count[i]; ++k) {
                int j = nbr_j(i, k);
                T w = kernel_w(length(xi - qv(j)), h, kind);
                s = s + w * get3(vin, j);
                den = den + w;
            }
            put3(vout, i, s / den);
        }
    }
    void shepard_surf(const std::vector<T>& kin, std::vector<T>& kout) {  // csf.py:175-194
        for (int i = 0; i < n; ++i) {
            kout[i] = kin[i];
            if (t_surf[i] == 0) continue;
            V3<T> xi = qv(i);
            T w0 = kernel_w(T(0), h, kind), num = w0 * kin[i], den = w0;
            for (int k = 0; k < count[i]; ++k) {
                int j = nbr_j(i, k);
                if (t_surf[j] == 0) continue;
                T w = kernel_w(length(xi - qv(j)), h, kind);
                num = num + w * kin[j];
                den = den + w;
            }
            kout[i] = num / den;
        }
    }
    void csf_stage(T dt) {  // csf.py:311-332
        for (int i = 0; i < n; ++i) {  // max_count_kernel, grad_kernel
            int s = count[i];
            for (int k = 0; k < count[i]; ++k) s = std::max(s, count[nbr_j(i, k)]);
            t_mcount[i] = T(s);
            V3<T> g{T(0), T(0), T(0)};
            if (is_fluid(i)) {
                V3<T> xi = qv(i);
                T inv = T(1) / rho0[i];
                for (int k = 0; k < count[i]; ++k) {
                    int j = nbr_j(i, k);
                    g = g + (mass[j] * inv) * kernel_grad(xi - qv(j), h, kind);
                }
                for (int b = 0; b < 4; ++b)
                    if (s_shape[4 * size_t(i) + b] >= 0) g = g + T(0.5) * get3(s_gpsi, 4 * size_t(i) + b);
            }
            put3(t_grad, i, g);
        }
        shepard_vec(t_grad, t_g1);
        shepard_vec(t_g1, t_g2);
        shepard_vec(t_g2, t_g1);
        for (int i = 0; i < n; ++i) {  // normal_kernel
            const V3<T> z{T(0), T(0), T(0)};
            put3(t_nhat, i, z);
            put3(t_nwall, i, z);
            t_wslot[i] = -1;
            t_surf[i] = 0;
            V3<T> gd = get3(t_g1, i);
            T lg = length(gd);
            if (!is_fluid(i) || length(get3(t_grad, i)) * h <= T(0.3) || lg <= T(0)) continue;
            V3<T> nn = (-gd) / lg;
            int best = -1;
            T dmin = T(1.0e30);
            for (int b = 0; b < 4; ++b)
                if (s_shape[4 * size_t(i) + b] >= 0 && s_phi[4 * size_t(i) + b] < dmin) {
                    dmin = s_phi[4 * size_t(i) + b];
                    best = b;
                }
            if (best >= 0) {
                int s = s_shape[4 * size_t(i) + best];
                if (s < S && sh_sin[s] >= T(0)) {
                    V3<T> nw = qv(i) - get3(s_xb, 4 * size_t(i) + best);
                    T lw = length(nw);
                    if (lw > T(1.0e-9)) {
                        nw = nw / lw;
                        if (dmin < T(0)) nw = -nw;
                        V3<T> t = nn - dot(nn, nw) * nw;
                        T lt = length(t);
                        if (lt > T(1.0e-6)) {
                            nn = sh_cos[s] * nw + (sh_sin[s] / lt) * t;
                            put3(t_nwall, i, nw);
                            t_wslot[i] = best;
                        }
                    }
                }
            }
            put3(t_nhat, i, nn);
            t_surf[i] = 1;
        }
        for (int i = 0; i < n; ++i) {  // curvature_kernel
            t_kraw[i] = T(0);
            if (t_surf[i] == 0) continue;
            V3<T> xi = qv(i), ni = get3(t_nhat, i);
            const T nv[3] = {ni.x, ni.y, ni.z};
            T P[9], A[9], B[9];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    P[3 * a + b] = (a == b ? T(1) : T(0)) - nv[a] * nv[b];
                    A[3 * a + b] = T(0);
                    B[3 * a + b] = T(0);
                }
            int cnt = 0;
            for (int k = 0; k < count[i]; ++k) {
                int j = nbr_j(i, k);
                if (t_surf[j] == 0) continue;
                V3<T> xj = qv(j);
                T w = kernel_w(length(xi - xj), h, kind);
                V3<T> t = mul33(P, xj - xi), d = mul33(P, get3(t_nhat, j) - ni);
                const T tv[3] = {t.x, t.y, t.z}, dv[3] = {d.x, d.y, d.z};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        A[3 * a + b] = A[3 * a + b] + w * (tv[a] * tv[b]);
                        B[3 * a + b] = B[3 * a + b] + w * (dv[a] * tv[b]);
                    }
                cnt += 1;
            }
            T scale = A[0] + A[4] + A[8];
            if (cnt < 4 || scale <= T(0)) continue;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) A[3 * a + b] = A[3 * a + b] + (T(1.0e-2) * scale) * (nv[a] * nv[b]);
            T Ai[9];
            inverse33(A, Ai);
            T tr = T(0);
            for (int a = 0; a < 3; ++a)
                tr = tr + (B[3 * a] * Ai[a] + B[3 * a + 1] * Ai[3 + a] + B[3 * a + 2] * Ai[6 + a]);
            t_kraw[i] = tr;
        }
        shepard_surf(t_kraw, t_ktmp);
        shepard_surf(t_ktmp, t_kraw);
        shepard_surf(t_kraw, t_ktmp);
        for (int i = 0; i < n; ++i) {  // csf_kernel
            t_kappa[i] = T(0);
            T wi = csf_weight(t_mcount[i], csf_lo, csf_hi);
            if (t_surf[i] == 0 || sigma[i] == T(0) || wi <= T(0)) continue;
            T kap = t_ktmp[i];
            T lg = length(get3(t_grad, i));
            T corr = T(1) + T(0.5) * kap * (ratio[i] - csf_cs) / lg;
            if (corr > T(0.5)) kap = kap / corr;
            t_kappa[i] = kap;
            T rho_i = tmax(ratio[i] * rho0[i], T(1.0e-6));
            V3<T> dv = (-(wi * dt * sigma[i] * kap * lg * csf_inv_i / rho_i)) * get3(t_nhat, i);
            int b = t_wslot[i];
            if (b >= 0) {
                V3<T> nw = get3(t_nwall, i);
                T dn = dot(dv, nw);
                if (dn > T(0)) {
                    dv = dv - dn * nw;
                    size_t sl = 4 * size_t(i) + b;
                    put3(s_imp, sl, get3(s_imp, sl) + (mass[i] * dn) * nw);
                }
            }
            setv(i, vv(i) + dv);
        }
        for (int i = 0; i < n; ++i) {  // normals_kernel (tension.py:101-115)
            V3<T> nn{T(0), T(0), T(0)};
            if (is_fluid(i)) {
                V3<T> xi = qv(i);
                for (int k = 0; k < count[i]; ++k) {
                    int j = nbr_j(i, k);
                    T rho_j = ratio[j] * rho0[j];
                    if (rho_j > T(0)) nn = nn + (mass[j] / rho_j) * kernel_grad(xi - qv(j), h, kind);
                }
                nn = h * nn;
            }
            put3(t_anorm, i, nn);
        }
        std::vector<T> dvp(3 * size_t(std::max(n, 1)), T(0));
        std::vector<char> upd(std::max(n, 1), 0);
        for (int i = 0; i < n; ++i) {  // sparse_pair_kernel (csf_sparse.py:65-88)
            if (!is_fluid(i) || gamma_[i] == T(0)) continue;
            V3<T> xi = qv(i), ni = get3(t_anorm, i), a{T(0), T(0), T(0)};
            T rho_i = ratio[i] * rho0[i];
            T wi = csf_weight(t_mcount[i], csf_lo, csf_hi);
            for (int k = 0; k < count[i]; ++k) {
                int j = nbr_j(i, k);
                T cc = T(1) - T(0.5) * (wi + csf_weight(t_mcount[j], csf_lo, csf_hi));
                T g = T(0.5) * (gamma_[i] + gamma_[j]);
                T rho_j = ratio[j] * rho0[j];
                if (cc <= T(0) || g == T(0) || rho_i + rho_j <= T(0)) continue;
                V3<T> rv = xi - qv(j);
                T r = length(rv);
                V3<T> acc = (T(4) * tmax(T(1) - r / h, T(0))) * (ni - get3(t_anorm, j));
                if (r > T(1.0e-6) * h) acc = acc + (mass[j] * cohesion_c(r) / r) * rv;
                a = a - (cc * T(2) * rho0[i] / (rho_i + rho_j) * g) * acc;
            }
            put3(dvp, i, dt * a);
            upd[i] = 1;
        }
        for (int i = 0; i < n; ++i)
            if (upd[i]) setv(i, vv(i) + get3(dvp, i));
    }

    // ---------------------------------------------------------------------------- shape meters (bodies.py:37-)
    void meters() {
        if (one_way) return;  // bodies.py:58-59
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 4; ++k) {
                int s = s_shape[4 * i + k];
                if (s < 0 || s >= S) continue;
                V3<T> J = get3(s_imp, 4 * i + k);
                if (J.x == T(0) && J.y == T(0) && J.z == T(0)) continue;
                const double sc = 1125899906842624.0;
                meter[3 * s] += int64_t(std::round(double(J.x) * sc));
                meter[3 * s + 1] += int64_t(std::round(double(J.y) * sc));
                meter[3 * s + 2] += int64_t(std::round(double(J.z) * sc));
            }
    }

    // ---------------------------------------------------------------------------- implicit viscosity (viscosity.py)
    T visc_coeff(int i, int j, V3<T> xij) const {  // viscosity.py:39-47
        T rho_j = ratio[j] * rho0[j];
        if (rho_j <= T(0)) return T(0);
        T mu_ij = T(0.5) * (mu[i] + mu[j]);
        return visc_d * mu_ij * (mass[j] / rho_j) / (dot(xij, xij) + T(0.01) * h * h);
    }
    static V3<T> mul33(const T* m, V3<T> v) {
        return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
                m[6] * v.x + m[7] * v.y + m[8] * v.z};
    }
    // the inverse by cofactors over the determinant (expanded along the first row); a singular block gives zeros
    static void inverse33(const T* m, T* r) {
        T c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
        T det = m[0] * c00 + m[1] * c01 + m[2] * c02;
        if (det == T(0)) {
            for (int k = 0; k < 9; ++k) r[k] = T(0);
            return;
        }
        T inv = T(1) / det;
        r[0] = c00 * inv;
        r[1] = (m[2] * m[7] - m[1] * m[8]) * inv;
        r[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
        r[3] = c01 * inv;
        r[4] = (m[0] * m[8] - m[2] * m[6]) * inv;
        r[5] = (m[2] * m[3] - m[0] * m[5]) * inv;
        r[6] = c02 * inv;
        r[7] = (m[1] * m[6] - m[0] * m[7]) * inv;
        r[8] = (m[0] * m[4] - m[1] * m[3]) * inv;
    }
    T rho_i_of(int i) const { return tmax(ratio[i] * rho0[i], T(1.0e-6)); }
    // wall_viscosity.py:96-105 _tangents
    static void wall_tangents(V3<T> nv, V3<T>& t1, V3<T>& t2) {
        V3<T> a{T(1), T(0), T(0)};
        if (std::abs(nv.x) > T(0.6)) a = V3<T>{T(0), T(1), T(0)};
        t1 = normalize(cross(nv, a));
        t2 = cross(nv, t1);
    }
    // wall_viscosity.py:108-150 slot_pair: M (row major) and w = M v_b of slot k of particle i; false when empty
    bool wall_slot_pair(int i, int k, V3<T> xi, V3<T> shift, T* m, V3<T>& w) const {
        size_t sl = 4 * size_t(i) + k;
        if (s_shape[sl] < 0) return false;
        T phi = s_phi[sl];
        T d = tmax(phi + T(0.5) * r, T(2) * r);
        T wd = kernel_w(d, h, kind);
        V3<T> nv = xi - get3(s_xb, sl);
        T ln = length(nv);
        T psi = s_psi[sl];
        if (phi > taper0) {
            T v0, v1;
            halfspace_volume(phi + offset, h, T(0), kind, v0, v1);
            psi = v0 * amplitude;
        }
        if (wd <= T(0) || psi <= T(0) || ln < T(1.0e-9)) return false;
        nv = nv / ln;
        V3<T> xk = xi - d * nv, t1, t2;
        wall_tangents(nv, t1, t2);
        T vol = T(0.25) * psi / wd;
        for (int q = 0; q < 9; ++q) m[q] = T(0);
        for (int q = 0; q < 4; ++q) {
            T sg = T(1) - T(2) * T(q % 2);
            V3<T> xis = xi - (xk + (sg * T(0.5) * h) * (q < 2 ? t1 : t2));
            V3<T> g = kernel_grad(xis, h, kind);
            T cc = wall_coeff * vol / (dot(xis, xis) + T(0.01) * h * h);
            const T gv[3] = {g.x, g.y, g.z}, xv[3] = {xis.x, xis.y, xis.z};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) m[3 * a + b] = m[3 * a + b] + cc * (gv[a] * xv[b]);
        }
        T u = tclamp((phi - T(2) * r) / r, T(0), T(1));
        T m_tt = -mu_b / (tmax(phi, r) * T(2) * r) * (T(1) - u * u * (T(3) - T(2) * u));
        T mn = dot(nv, mul33(m, nv));
        const T nn3[3] = {nv.x, nv.y, nv.z};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                T nn = nn3[a] * nn3[b];
                m[3 * a + b] = mn * nn + m_tt * ((a == b ? T(1) : T(0)) - nn);
            }
        w = mul33(m, get3(s_vb, sl) + shift);
        return true;
    }
    // solver.py:310-324 _wall_terms with viscosity.py:235-323 (the pressure guess's shift) and
    // wall_viscosity.py:153-167 (wall_terms_kernel)
    void visc_wall_terms(T dt, bool predict) {
        if (predict) {
            for (int i = 0; i < n; ++i) vgk[i] = is_fluid(i) ? tmin(kd[i], T(2.5e-4)) / (dt * dt) : T(0);
            for (int i = 0; i < n; ++i) {
                if (!is_fluid(i)) {
                    put3(vraw, i, V3<T>{T(0), T(0), T(0)});
                    continue;
                }
                V3<T> xi = qv(i), a{T(0), T(0), T(0)};
                T ki = vgk[i] / rho0[i];
                for (int k = 0; k < count[i]; ++k) {
                    int j = nbr_j(i, k);
                    V3<T> g = kernel_grad(xi - qv(j), h, kind);
                    a = a - (mass[j] * (ki + vgk[j] / rho0[j])) * g;
                }
                for (int b = 0; b < 4; ++b)
                    if (s_shape[4 * size_t(i) + b] >= 0) a = a - vgk[i] * get3(s_gpsi, 4 * size_t(i) + b);
                put3(vraw, i, (-dt) * a);
            }
            for (int i = 0; i < n; ++i) {
                bool wall = false;
                for (int b = 0; b < 4; ++b) wall = wall || s_shape[4 * size_t(i) + b] >= 0;
                if (!is_fluid(i) || !wall) {
                    put3(vshift, i, V3<T>{T(0), T(0), T(0)});
                    continue;
                }
                V3<T> xi = qv(i);
                T tot = mass[i] / rho0[i] * kernel_w(T(0), h, kind);
                V3<T> acc_ = tot * get3(vraw, i);
                for (int k = 0; k < count[i]; ++k) {
                    int j = nbr_j(i, k);
                    if (!is_fluid(j)) continue;
                    T wj = mass[j] / rho0[j] * kernel_w(length(xi - qv(j)), h, kind);
                    acc_ = acc_ + wj * get3(vraw, j);
                    tot = tot + wj;
                }
                put3(vshift, i, acc_ / tot);
            }
        }
        for (int i = 0; i < n; ++i) {
            T m[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
            V3<T> w{T(0), T(0), T(0)};
            if (is_fluid(i)) {
                V3<T> xi = qv(i), sh = predict ? get3(vshift, i) : V3<T>{T(0), T(0), T(0)};
                for (int b = 0; b < 4; ++b) {
                    T mk[9];
                    V3<T> wk;
                    if (!wall_slot_pair(i, b, xi, sh, mk, wk)) continue;
                    for (int q = 0; q < 9; ++q) m[q] = m[q] + mk[q];
                    w = w + wk;
                }
            }
            for (int q = 0; q < 9; ++q) wall_m[9 * size_t(i) + q] = m[q];
            put3(wall_w, i, w);
        }
    }
    // wall_viscosity.py:170-184 wall_reaction_kernel
    void visc_wall_reaction(T dt, bool predict) {
        for (int i = 0; i < n; ++i) {
            if (!is_fluid(i)) continue;
            V3<T> xi = qv(i), vi = vv(i), sh = predict ? get3(vshift, i) : V3<T>{T(0), T(0), T(0)};
            T f = mass[i] * dt / rho_i_of(i);
            for (int b = 0; b < 4; ++b) {
                T mk[9];
                V3<T> wk;
                if (!wall_slot_pair(i, b, xi, sh, mk, wk)) continue;
                size_t sl = 4 * size_t(i) + b;
                put3(s_imp, sl, get3(s_imp, sl) - f * (mul33(mk, vi) - wk));
            }
        }
    }
    // viscosity.py:50-71: out = A vec, vals[2i] = vec_i . out_i (converged worlds untouched)
    void visc_apply(T dt, const std::vector<T>& vec, std::vector<T>& out, std::vector<T>& vals) {
        for (int i = 0; i < n; ++i) {
            vals[2 * i] = vals[2 * i + 1] = T(0);
            if (v_conv[world[i] < 0 ? 0 : world[i]] != 0 && world[i] >= 0) continue;
            if (!is_fluid(i)) {
                put3(out, i, V3<T>{T(0), T(0), T(0)});
                continue;
            }
            V3<T> xi = qv(i), vi = get3(vec, i), a{T(0), T(0), T(0)};
            for (int k = 0; k < count[i]; ++k) {
                int j = nbr_j(i, k);
                V3<T> xij = xi - qv(j);
                a = a + (visc_coeff(i, j, xij) * dot(xij, vi - get3(vec, j))) * kernel_grad(xij, h, kind);
            }
            V3<T> r = vi - (dt / rho_i_of(i)) * a;
            if (wall_coeff > T(0)) r = r - (dt / rho_i_of(i)) * mul33(&wall_m[9 * size_t(i)], vi);
            put3(out, i, r);
            vals[2 * i] = dot(vi, r);
        }
    }
    void visc_solve(T dt, bool predict) {  // viscosity.py:243-283
        std::vector<T> vals(2 * size_t(std::max(n, 1)), T(0));
        std::vector<double> sums(2 * size_t(W));
        std::vector<T> mx(W);
        if (wall_coeff > T(0)) visc_wall_terms(dt, predict);
        for (int i = 0; i < n; ++i) {
            put3(vb, i, vv(i));
            put3(vx, i, vv(i));
        }
        std::fill(v_conv.begin(), v_conv.end(), 0);
        visc_apply(dt, vx, vap, vals);
        for (int i = 0; i < n; ++i) {  // visc_init_kernel
            vals[2 * i] = vals[2 * i + 1] = T(0);
            if (!is_fluid(i)) {
                for (auto* v : {&vr, &vz, &vp}) put3(*v, i, V3<T>{T(0), T(0), T(0)});
                continue;
            }
            V3<T> xi = qv(i);
            T rho_i = rho_i_of(i);
            T blk[9] = {T(1), T(0), T(0), T(0), T(1), T(0), T(0), T(0), T(1)};
            for (int k = 0; k < count[i]; ++k) {
                int j = nbr_j(i, k);
                V3<T> xij = xi - qv(j);
                T s = (dt / rho_i) * visc_coeff(i, j, xij);
                V3<T> g = kernel_grad(xij, h, kind);
                const T gv[3] = {g.x, g.y, g.z}, xv[3] = {xij.x, xij.y, xij.z};
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) blk[3 * r + q] = blk[3 * r + q] - s * (gv[r] * xv[q]);
            }
            V3<T> bi = get3(vb, i);
            if (wall_coeff > T(0)) {
                T s = dt / rho_i;
                for (int q = 0; q < 9; ++q) blk[q] = blk[q] - s * wall_m[9 * size_t(i) + q];
                bi = bi - s * get3(wall_w, i);
                put3(vb, i, bi);
            }
            inverse33(blk, &vminv[9 * size_t(i)]);
            V3<T> ri = bi - get3(vap, i);
            V3<T> zi = mul33(&vminv[9 * size_t(i)], ri);
            put3(vr, i, ri);
            put3(vz, i, zi);
            put3(vp, i, zi);
            vals[2 * i] = dot(ri, zi);
            vals[2 * i + 1] = dot(bi, bi);
        }
        reduce(vals, 2, sums, mx.data());
        int unconv = 0;
        for (int w = 0; w < W; ++w) {  // visc_begin_kernel
            v_rz[w] = sums[2 * w];
            v_bb[w] = sums[2 * w + 1];
            v_iters[w] = 0;
            v_conv[w] = (n_active[w] > 0.0 && sums[2 * w] > 0.0) ? 0 : 1;
            unconv += v_conv[w] == 0;
        }
        while (unconv > 0) {
            visc_apply(dt, vp, vap, vals);
            reduce(vals, 2, sums, mx.data());
            for (int w = 0; w < W; ++w) {  // visc_alpha_kernel
                if (v_conv[w]) continue;
                double a = 0.0;
                if (sums[2 * w] > 1.0e-30) a = v_rz[w] / sums[2 * w];
                v_step[w] = a;
            }
            for (int i = 0; i < n; ++i) {  // visc_update_kernel
                vals[2 * i] = vals[2 * i + 1] = T(0);
                if (!is_fluid(i) || v_conv[world[i]]) continue;
                T a = T(v_step[world[i]]);
                put3(vx, i, get3(vx, i) + a * get3(vp, i));
                V3<T> ri = get3(vr, i) - a * get3(vap, i);
                V3<T> zi = mul33(&vminv[9 * size_t(i)], ri);
                put3(vr, i, ri);
                put3(vz, i, zi);
                vals[2 * i] = dot(ri, zi);
                vals[2 * i + 1] = dot(ri, ri);
            }
            reduce(vals, 2, sums, mx.data());
            unconv = 0;
            for (int w = 0; w < W; ++w) {  // visc_check_kernel
                if (v_conv[w]) continue;
                int it = v_iters[w] + 1;
                v_iters[w] = it;
                double beta = 0.0;
                if (v_rz[w] > 1.0e-30) beta = sums[2 * w] / v_rz[w];
                v_rz[w] = sums[2 * w];
                v_step[w] = beta;
                if (sums[2 * w + 1] <= double(visc_tol * visc_tol) * v_bb[w] || it >= visc_max_it) v_conv[w] = 1;
                else unconv += 1;
            }
            for (int i = 0; i < n; ++i)  // visc_direction_kernel
                if (is_fluid(i) && !v_conv[world[i]]) put3(vp, i, get3(vz, i) + T(v_step[world[i]]) * get3(vp, i));
        }
        for (int i = 0; i < n; ++i)  // visc_store_kernel
            if (is_fluid(i)) setv(i, get3(vx, i));
        if (wall_coeff > T(0)) visc_wall_reaction(dt, predict);
    }

    // particle_walls.py:138-177 _accumulate_kernel (impulse split over the wall particles, the walls' work)
    void pw_accumulate() {
        std::fill(pw_work.begin(), pw_work.end(), 0);
        if (NW == 0) return;
        const double IMP = 1125899906842624.0;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 4; ++k) {
                if (s_shape[4 * i + k] != S) continue;
                V3<T> J = get3(s_imp, 4 * size_t(i) + k);
                if (J.x == T(0) && J.y == T(0) && J.z == T(0)) continue;
                int w = world[i];
                pw_work[w] += std::llround(double(-dot(J, get3(s_vb, 4 * size_t(i) + k))) * IMP);
                if (one_way) continue;
                V3<T> x = qv(i), gc, mv;
                T col;
                pw_sample(x, w, col, gc, mv);
                if (col <= T(0)) continue;
                int64_t left[3] = {std::llround(double(J.x) * IMP), std::llround(double(J.y) * IMP),
                                   std::llround(double(J.z) * IMP)};
                int best = -1;
                T sbest = T(0);
                int g = W + w, cx, cy, cz;
                cell_of(g, x, cx, cy, cz);
                const int* d = &g_dims[3 * w];
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            int X = cx + dx, Y = cy + dy, Z = cz + dz;
                            if (X < 0 || Y < 0 || Z < 0 || X >= d[0] || Y >= d[1] || Z >= d[2]) continue;
                            int cc = cell_base[g] + (X * d[1] + Y) * d[2] + Z;
                            for (int e = cell_start[cc]; e < cell_start[cc + 1]; ++e) {
                                int j = cell_items[e];
                                T sj = wvol[j] * kernel_w(length(x - qv(j)), h, kind) / col;
                                if (sj > T(0)) {
                                    int cj = wcidx[j];
                                    int64_t part[3] = {std::llround(double(J.x * sj) * IMP),
                                                       std::llround(double(J.y * sj) * IMP),
                                                       std::llround(double(J.z * sj) * IMP)};
                                    for (int q = 0; q < 3; ++q) {
                                        pw_exch[3 * size_t(cj) + q] += part[q];
                                        left[q] -= part[q];
                                    }
                                    if (sj > sbest) {
                                        sbest = sj;
                                        best = cj;
                                    }
                                }
                            }
                        }
                if (best >= 0)
                    for (int q = 0; q < 3; ++q) pw_exch[3 * size_t(best) + q] += left[q];
            }
    }

    // ---------------------------------------------------------------------------- advect / project (integrate.py)
    T path_sdf(int s, V3<T> l0, V3<T> l1, T t) const {
        T phi;
        V3<T> g;
        shape_sdf(sh_geo[s], sh_scale[s], l0 + t * (l1 - l0), phi, g);
        return phi;
    }
    void box_span(V3<T> half, V3<T> l0, V3<T> l1, T& lo, T& hi) const {
        lo = T(-1.0e30);
        hi = T(1.0e30);
        V3<T> d = l1 - l0;
        T hv[3] = {half.x, half.y, half.z}, a0[3] = {l0.x, l0.y, l0.z}, dv[3] = {d.x, d.y, d.z};
        for (int a = 0; a < 3; ++a) {
            if (std::fabs(dv[a]) > T(0)) {
                T ta = (-hv[a] - a0[a]) / dv[a];
                T tb = (hv[a] - a0[a]) / dv[a];
                lo = tmax(lo, tmin(ta, tb));
                hi = tmin(hi, tmax(ta, tb));
            } else if (std::fabs(a0[a]) > hv[a]) {
                lo = T(2.0e30);
            }
        }
    }
    T path_entry(int s, V3<T> l0, V3<T> l1, int samples) const {
        if (path_sdf(s, l0, l1, T(0)) < T(0)) return T(2);
        if (sh_geo[s] == GEO_BOX) {
            T lo, hi;
            box_span(sh_scale[s], l0, l1, lo, hi);
            if (lo <= hi && lo <= T(1) && hi >= T(0)) return tmax(lo, T(0));
            return T(2);
        }
        T t_out = T(2), t_lo = T(0), t_hi = T(2);
        for (int k = 1; k <= samples; ++k) {
            T t = T(k) / T(samples);
            if (t_hi > T(1)) {
                if (path_sdf(s, l0, l1, t) < T(0)) t_hi = t;
                else t_lo = t;
            }
        }
        if (t_hi <= T(1)) {
            for (int b = 0; b < 8; ++b) {
                T tm = T(0.5) * (t_lo + t_hi);
                if (path_sdf(s, l0, l1, tm) < T(0)) t_hi = tm;
                else t_lo = tm;
            }
            t_out = t_lo;
        }
        return t_out;
    }
    int sweep_entry(int w, V3<T> x0, V3<T> x1, V3<T>& l_hit) const {
        T len = length(x1 - x0);
        T reach_ = len + r;  // spacing argument of sweep_entry is the particle radius (solver.py:258-261)
        int samples = std::min(std::max(int(std::ceil(len / r)) + 1, 8), 64);
        T best = T(2);
        int hit = -1;
        l_hit = v3<T>(0, 0, 0);
        for (int c = sh_start[w]; c < sh_start[w + 1]; ++c) {
            int s = sh_ids[c];
            const Xform<T>& X = sh_x[s];
            bool near_ = sh_geo[s] == GEO_PLANE || length(x0 - X.p) <= sh_radius[s] + reach_;
            if (!near_) continue;
            V3<T> l0 = xf_inv_point(X, x0), l1 = xf_inv_point(X, x1);
            T t = path_entry(s, l0, l1, samples);
            if (t < best) {
                best = t;
                hit = s;
                l_hit = l0 + t * (l1 - l0);
            }
        }
        return hit;
    }
    V3<T> wall_friction(V3<T> v, V3<T> vb, V3<T> nn, T vn, int s) const {
        if (friction == 0) return v;
        T mu = sh_mu[s];
        if (mu <= T(0)) return v;
        V3<T> u = v - vb;
        V3<T> vt = u - dot(u, nn) * nn;
        T lt = length(vt);
        if (lt <= T(0)) return v;
        return v - tmin(T(1), mu * (-vn) / lt) * vt;
    }

    void advect_project(T dt) {
        for (int i = 0; i < n; ++i)  // advect_kernel
            if (is_fluid(i)) setq(i, qv(i) + dt * vv(i));
        for (int j = 0; j < n; ++j)  // particle walls advance with their own velocity (particle_walls.py:131-135)
            if (NW > 0 && wvol[j] > T(0) && !is_fluid(j)) setq(j, qv(j) + dt * vv(j));
        for (int i = 0; i < n; ++i) {  // project_kernel
            if (!is_fluid(i)) continue;
            V3<T> x = qv(i), v = vv(i);
            int moved = 0;
            if (length(dt * v) > sweep_len) {
                V3<T> l;
                int s = sweep_entry(world[i], x - dt * v, x, l);
                if (s >= 0) {
                    const Xform<T>& X = sh_x[s];
                    T phi;
                    V3<T> gl;
                    shape_sdf(sh_geo[s], sh_scale[s], l, phi, gl);
                    V3<T> nn = normalize(xf_vector(X, gl));
                    x = xf_point(X, l) + tmax(eps - phi, T(0)) * nn;
                    V3<T> vb{0, 0, 0};
                    T vn = dot(v - vb, nn);
                    if (vn < T(0)) {
                        v = v - vn * nn;
                        v = wall_friction(v, vb, nn, vn, s);
                    }
                    moved = 1;
                    st_swept[world[i]] += 1;
                }
            }
            int busy = 1, npass = 0;
            while (busy != 0 && npass < 4) {
                busy = 0;
                npass += 1;
                for (int b = 0; b < 4; ++b) {
                    size_t sk = 4 * size_t(i) + b;
                    int s = s_shape[sk];
                    if (s == S && NW > 0) {  // pw_project (particle_walls.py:122-128)
                        V3<T> vb = get3(s_vb, sk), nb = get3(s_nbl, sk), xb = get3(s_xbl, sk);
                        T d = dot(x - (xb + dt * vb), nb);
                        if (d < eps) {
                            x = x + (eps - d) * nb;
                            T vn = dot(v - vb, nb);
                            if (vn < T(0)) v = v - vn * nb;
                            moved = 1;
                            busy = 1;
                        }
                    } else if (s >= 0 && s < S) {
                        const Xform<T>& X = sh_x[s];
                        V3<T> xl = xf_inv_point(X, x);
                        T phi;
                        V3<T> gl;
                        shape_sdf(sh_geo[s], sh_scale[s], xl, phi, gl);
                        if (phi < eps) {
                            V3<T> nn = normalize(xf_vector(X, gl));
                            T push = eps - phi;
                            if (s_phi[sk] >= T(0)) {
                                V3<T> ns = xf_vector(X, get3(s_nbl, sk));
                                T push_s = eps - dot(ns, x - xf_point(X, get3(s_xbl, sk)));
                                if (push_s > T(0)) {
                                    nn = ns;
                                    push = push_s;
                                }
                            }
                            if (push > T(0)) {
                                x = x + push * nn;
                                V3<T> vb{0, 0, 0};
                                T vn = dot(v - vb, nn);
                                if (vn < T(0)) {
                                    v = v - vn * nn;
                                    v = wall_friction(v, vb, nn, vn, s);
                                }
                                moved = 1;
                                busy = 1;
                            }
                        }
                    }
                }
            }
            if (moved) {
                setq(i, x);
                setv(i, v);
                st_projected[world[i]] += 1;
            }
        }
        for (int w = 0; w < W; ++w) st_max_speed[w] = T(0);
        if (!emit_active) speed();  // with emitters, sinks or an escape box, after their pass
    }

    // speed_kernel (integrate.py:312-326)
    void speed() {
        for (int i = 0; i < n; ++i) {
            if (!is_fluid(i)) continue;
            V3<T> v = vv(i), x = qv(i);
            T s = length(v);
            int w = world[i];
            if (std::isfinite(s) && std::isfinite(x.x + x.y + x.z)) st_max_speed[w] = tmax(st_max_speed[w], s);
            else {
                st_nonfinite[w] += 1;
                st_max_speed[w] = tmax(st_max_speed[w], T(1.0e30));
            }
        }
    }

    // ---------------------------------------------------------------------------- emitters
    static int64_t fixed(double x, double scale) { return std::llround(x * scale); }  // bodies.py:31-32 _fx
    static int64_t to_fixed(T m) { return fixed(double(m), 1099511627776.0); }      // kernels.py:142-144

    // emitters.py:82-89 flux_add
    void flux_add(int w, T m, V3<T> v, V3<T> x, T sign) {
        V3<T> g = get3(gravity, w);
        T en = sign * m * (T(0.5) * dot(v, v) - dot(g, x));
        flux_e[w] += fixed(double(en), 1125899906842624.0);
        flux_g[w] += fixed(double(sign * m * g.z), 1099511627776.0);
    }

    // emitters.py:127-130 _in_box
    static bool in_box(const Xform<T>& X, V3<T> half, V3<T> x) {
        V3<T> xl = xf_inv_point(X, x);
        return std::fabs(xl.x) <= half.x && std::fabs(xl.y) <= half.y && std::fabs(xl.z) <= half.z;
    }

    // emitters.py:133-165 retire_kernel
    void retire() {
        for (int i = 0; i < n; ++i) {
            if (!is_fluid(i)) continue;
            int w = world[i];
            V3<T> x = qv(i);
            int gone = 0;
            for (int k = sink_start[w]; k < sink_start[w + 1]; ++k)
                if (in_box(s_xform[k], s_half[k], x)) gone = 1;
            if (has_bounds && gone == 0) {
                bool out = x.x < lo.x || x.y < lo.y || x.z < lo.z || x.x > hi.x || x.y > hi.y || x.z > hi.z;
                if (out || !std::isfinite(x.x + x.y + x.z)) {
                    gone = 2;
                    escaped[w] += 1;
                }
            }
            if (gone != 0) {
                V3<T> v = vv(i);
                if (std::isfinite(x.x + x.y + x.z + v.x + v.y + v.z)) flux_add(w, mass[i], v, x, T(-1));
                flags[i] = flags[i] & ~ACTIVE;
                phase[i] = int8_t(-2);
                setv(i, V3<T>{T(0), T(0), T(0)});
                mass_out[w] += to_fixed(mass[i]);
            }
        }
    }

    // emitters.py:67-79 emit_clock_kernel, 92-117 emit_kernel, 120-124 advance_cursor_kernel
    void emit(T dt) {
        std::vector<int> now(E, 0);
        for (int e = 0; e < E; ++e) {
            if (enabled[e] == 0) continue;
            T t = t_acc[e] + dt;
            if (t >= e_interval[e]) {
                t -= e_interval[e];
                now[e] = 1;
            }
            t_acc[e] = t;
        }
        for (int e = 0; e < E; ++e) {
            if (!now[e]) continue;
            int w = e_world[e], psize = pool_start[e + 1] - pool_start[e];
            for (int r = 0; r < off_start[e + 1] - off_start[e]; ++r) {
                int slot = cursor[e] + r;
                if (slot >= psize) {
                    emit_overflow[w] += 1;
                    continue;
                }
                int i = pool_ids[pool_start[e] + slot];
                setq(i, xf_point(e_xform[e], off_local[off_start[e] + r]));
                setv(i, e_vel[e]);
                flags[i] = flags[i] | ACTIVE;
                phase[i] = int8_t(1);
                mass_in[w] += to_fixed(mass[i]);
                flux_add(w, mass[i], vv(i), qv(i), T(1));
            }
        }
        for (int e = 0; e < E; ++e)
            if (now[e]) cursor[e] = std::min(cursor[e] + (off_start[e + 1] - off_start[e]), pool_start[e + 1] - pool_start[e]);
    }

    // runtime.py:318-326 EmitterManager.step, then the step's speed statistics (solver.py:266-267)
    void emitters_step(T dt) {
        if (!emit_active) return;
        retire();
        if (E > 0) emit(dt);
        speed();
    }

    // runtime.py:126-130 _active_mass
    void active_mass() {
        std::fill(mass_act.begin(), mass_act.end(), 0);
        for (int i = 0; i < n; ++i)
            if (is_fluid(i)) mass_act[world[i]] += to_fixed(mass[i]);
    }

    int set_emitters(const DfsphEmitters* t) {
        int En = t->emitter_count, K = t->sink_count;
        for (int e = 0; e < En; ++e) {
            int w = t->world[e];
            if (w < 0 || w >= W) return fail(DFSPH_EINVAL, "add_emitter: world out of range");
            for (int k = t->pool_start[e]; k < t->pool_start[e + 1]; ++k)
                if (t->pool_ids[k] < ws[w] || t->pool_ids[k] >= ws[w + 1])
                    return fail(DFSPH_EINVAL, "pool id is not a particle of the emitter's world");
        }
        for (int k = 0; k < K; ++k)
            if (t->sink_world[k] < 0 || t->sink_world[k] >= W) return fail(DFSPH_EINVAL, "add_sink: world out of range");
        std::vector<T> t_old = t_acc;
        std::vector<int> c_old = cursor, en_old = enabled;
        E = En;
        e_world.assign(t->world, t->world + En);
        e_xform.resize(En);
        e_vel.resize(En);
        e_interval.resize(En);
        for (int e = 0; e < En; ++e) {
            const float* x = t->xform + 7 * e;
            e_xform[e] = Xform<T>{V3<T>{T(x[0]), T(x[1]), T(x[2])}, T(x[3]), T(x[4]), T(x[5]), T(x[6])};
            e_vel[e] = V3<T>{T(t->velocity[3 * e]), T(t->velocity[3 * e + 1]), T(t->velocity[3 * e + 2])};
            e_interval[e] = T(t->interval[e]);
        }
        off_start.assign(t->offset_start, t->offset_start + (En > 0 ? En + 1 : 0));
        pool_start.assign(t->pool_start, t->pool_start + (En > 0 ? En + 1 : 0));
        int n_off = En > 0 ? off_start[En] : 0, n_pool = En > 0 ? pool_start[En] : 0;
        off_local.resize(n_off);
        for (int k = 0; k < n_off; ++k)
            off_local[k] = V3<T>{T(t->offsets[3 * k]), T(t->offsets[3 * k + 1]), T(t->offsets[3 * k + 2])};
        pool_ids.assign(t->pool_ids, t->pool_ids + n_pool);
        t_acc.assign(En, T(0));
        cursor.assign(En, 0);
        enabled.assign(En, 1);
        for (int e = 0; e < std::min<int>(En, int(t_old.size())); ++e) {  // EmitterDevice.carry_from
            t_acc[e] = t_old[e];
            cursor[e] = c_old[e];
            enabled[e] = en_old[e];
        }
        sink_start.assign(W + 1, 0);
        for (int k = 0; k < K; ++k) sink_start[t->sink_world[k] + 1] += 1;
        for (int w = 0; w < W; ++w) sink_start[w + 1] += sink_start[w];
        s_xform.resize(K);
        s_half.resize(K);
        std::vector<int> fillp(sink_start.begin(), sink_start.end() - 1);
        for (int k = 0; k < K; ++k) {
            int d = fillp[t->sink_world[k]]++;
            const float* x = t->sink_xform + 7 * k;
            s_xform[d] = Xform<T>{V3<T>{T(x[0]), T(x[1]), T(x[2])}, T(x[3]), T(x[4]), T(x[5]), T(x[6])};
            s_half[d] = V3<T>{T(t->sink_half[3 * k]), T(t->sink_half[3 * k + 1]), T(t->sink_half[3 * k + 2])};
        }
        has_bounds = t->has_bounds != 0;
        lo = V3<T>{T(t->bounds_lo[0]), T(t->bounds_lo[1]), T(t->bounds_lo[2])};
        hi = V3<T>{T(t->bounds_hi[0]), T(t->bounds_hi[1]), T(t->bounds_hi[2])};
        emit_active = E > 0 || K > 0 || has_bounds;
        return DFSPH_OK;
    }

    // ---------------------------------------------------------------------------- prime (prime.py:53-115)
    int prime(const T* guess, double dt_host, double tol, int max_iterations) {
        T dt = T(dt_host);
        std::vector<T> qd_saved(qd, qd + 3 * size_t(n));
        build_neighbors();
        boundary_slots();
        density_alpha();
        for (int i = 0; i < n; ++i) kv[i] = guess[i] * dt;
        double s_tol = tol_div, s_fac = err_factor;
        int s_max = max_it;
        T s_warm = warm_div, s_nb = min_nb;
        tol_div = tol;
        max_it = max_iterations;
        warm_div = T(1);
        min_nb = T(0);
        err_factor = 1.0e9;
        nonpressure(dt);
        if (visc_on) visc_solve(dt, false);
        solve(0, dt_host, -1);
        tol_div = s_tol;
        err_factor = s_fac;
        max_it = s_max;
        warm_div = s_warm;
        min_nb = s_nb;
        std::copy(qd_saved.begin(), qd_saved.end(), qd);
        std::fill(s_imp.begin(), s_imp.end(), T(0));
        for (int i = 0; i < n; ++i) kd[i] = dt * kv[i];
        std::fill(kv.begin(), kv.end(), T(0));
        int m = 0;
        for (int w = 0; w < W; ++w) m = std::max(m, it_div[w]);
        return m;
    }

    // ---------------------------------------------------------------------------- step (solver.py:218-268)
    void step(double dt_host, const T* warm_div_k = nullptr, const T* warm_den_k = nullptr) {
        T dt = T(dt_host);
        std::fill(s_imp.begin(), s_imp.end(), T(0));
        build_neighbors();
        boundary_slots();
        density_alpha();
        flow_speed();
        if (div_solve) solve(0, dt_host, -1, warm_div_k);
        nonpressure(dt);
        if (visc_on) visc_solve(dt, true);  // solver.py:306-308
        solve(1, dt_host, -1, warm_den_k);
        meters();
        pw_accumulate();
        advect_project(dt);
        emitters_step(dt);
    }
};

struct Handle {
    int precision;
    void* impl;
};

template <class T> Oracle<T>* impl(void* h) { return static_cast<Oracle<T>*>(static_cast<Handle*>(h)->impl); }

template <class T>
int get_buffer_t(Oracle<T>* o, int which, void** ptr, int64_t* nb) {
#define OB(vec) \
    { *ptr = (void*)o->vec.data(); *nb = int64_t(o->vec.size() * sizeof(o->vec[0])); return DFSPH_OK; }
    switch (which) {
        case DFSPH_BUF_NBR: OB(nbr)
        case DFSPH_BUF_NBR_HI: OB(nbr_hi)
        case DFSPH_BUF_COUNT: OB(count)
        case DFSPH_BUF_GROUP: OB(group)
        case DFSPH_BUF_SLOT_SHAPE: OB(s_shape)
        case DFSPH_BUF_SLOT_PSI: OB(s_psi)
        case DFSPH_BUF_SLOT_GPSI: OB(s_gpsi)
        case DFSPH_BUF_SLOT_XB: OB(s_xb)
        case DFSPH_BUF_SLOT_VB: OB(s_vb)
        case DFSPH_BUF_SLOT_PHI: OB(s_phi)
        case DFSPH_BUF_SLOT_XBL: OB(s_xbl)
        case DFSPH_BUF_SLOT_NBL: OB(s_nbl)
        case DFSPH_BUF_SLOT_IMPULSE: OB(s_imp)
        case DFSPH_BUF_RATIO: OB(ratio)
        case DFSPH_BUF_ALPHA: OB(alpha)
        case DFSPH_BUF_NEFF: OB(neff)
        case DFSPH_BUF_ACTIVE: OB(active)
        case DFSPH_BUF_REG: OB(reg)
        case DFSPH_BUF_SRC: OB(src)
        case DFSPH_BUF_KAPPA: OB(kappa)
        case DFSPH_BUF_ACC: OB(acc)
        case DFSPH_BUF_ERR: OB(err)
        case DFSPH_BUF_LOWER: OB(lower)
        case DFSPH_BUF_KV: OB(kv)
        case DFSPH_BUF_KD: OB(kd)
        case DFSPH_BUF_VTMP: OB(vtmp)
        case DFSPH_BUF_CONVERGED: OB(converged)
        case DFSPH_BUF_ITERS: OB(iters)
        case DFSPH_BUF_ERR_MEAN: OB(err_mean)
        case DFSPH_BUF_ERR_MAX: OB(err_max)
        case DFSPH_BUF_N_ACTIVE: OB(n_active)
        case DFSPH_BUF_TOL_SCALE: OB(tol_scale)
        case DFSPH_BUF_OVERFLOW: OB(st_overflow)
        case DFSPH_BUF_MAX_SPEED: OB(st_max_speed)
        case DFSPH_BUF_ITERS_DEN: OB(it_den)
        case DFSPH_BUF_ITERS_DIV: OB(it_div)
        case DFSPH_BUF_DEN_ERR: OB(st_den_err)
        case DFSPH_BUF_DIV_ERR: OB(st_div_err)
        case DFSPH_BUF_SHAPE_METER: OB(meter)
        case DFSPH_BUF_MASS_INITIAL: OB(mass0)
        case DFSPH_BUF_MASS_IN: OB(mass_in)
        case DFSPH_BUF_MASS_OUT: OB(mass_out)
        case DFSPH_BUF_MASS_ACTIVE: OB(mass_act)
        case DFSPH_BUF_ESCAPED: OB(escaped)
        case DFSPH_BUF_EMIT_OVERFLOW: OB(emit_overflow)
        case DFSPH_BUF_FLUX_E: OB(flux_e)
        case DFSPH_BUF_FLUX_G: OB(flux_g)
        case DFSPH_BUF_EMIT_T: OB(t_acc)
        case DFSPH_BUF_EMIT_CURSOR: OB(cursor)
        case DFSPH_BUF_EMIT_ENABLED: OB(enabled)
        case DFSPH_BUF_FLAGS: OB(flags)
        case DFSPH_BUF_PHASE: OB(phase)
        case DFSPH_BUF_PW_EXCH: OB(pw_exch)
        case DFSPH_BUF_PW_WORK: OB(pw_work)
        case DFSPH_BUF_VISC_ITERS: OB(v_iters)
        case DFSPH_BUF_CSF_SURF: OB(t_surf)
        case DFSPH_BUF_CSF_KAPPA: OB(t_kappa)
        default: return fail(DFSPH_EINVAL, "unknown buffer id");
    }
#undef OB
}

template <class T>
void bind(Oracle<T>* o, void* q, void* qd, const void* f) {
    o->q = static_cast<T*>(q);
    o->qd = static_cast<T*>(qd);
    o->f = static_cast<const T*>(f);
}

}  // namespace

#define DISPATCH(h, call)                                                    \
    do {                                                                     \
        if (!(h)) return fail(DFSPH_EINVAL, "null handle");                  \
        if (static_cast<Handle*>(h)->precision == 64) {                      \
            auto* o = impl<double>(h);                                       \
            call;                                                            \
        } else {                                                             \
            auto* o = impl<float>(h);                                        \
            call;                                                            \
        }                                                                    \
    } while (0)

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// precision 32 or 64; model/layout arrays are host arrays
int32_t orc_create(int32_t precision, const DfsphParams* P, const DfsphLayout* L, const DfsphModel* M, void** out) {
    if (!P || !L || !M || !out) return fail(DFSPH_EINVAL, "null argument");
    Handle* h = new Handle{precision, nullptr};
    int rc;
    if (precision == 64) {
        auto* o = new Oracle<double>();
        h->impl = o;
        rc = o->init(P, L, M);
    } else {
        auto* o = new Oracle<float>();
        h->impl = o;
        rc = o->init(P, L, M);
    }
    if (rc != DFSPH_OK) {
        if (precision == 64) delete impl<double>(h);
        else delete impl<float>(h);
        delete h;
        return rc;
    }
    *out = h;
    return DFSPH_OK;
}

int32_t orc_destroy(void* h) {
    if (!h) return DFSPH_OK;
    if (static_cast<Handle*>(h)->precision == 64) delete impl<double>(h);
    else delete impl<float>(h);
    delete static_cast<Handle*>(h);
    return DFSPH_OK;
}

// warm_div / warm_den: optional [n] initial kappa of the two solves (learned warm start), NULL = reference rule
int32_t orc_step(void* h, void* q, void* qd, const void* f, double dt, const void* warm_div, const void* warm_den) {
    if (!(dt > 0.0)) return fail(DFSPH_EINVAL, "dt must be > 0");
    DISPATCH(h, {
        using TT = std::remove_pointer_t<decltype(o->q)>;
        bind(o, q, qd, f);
        o->step(dt, static_cast<const TT*>(warm_div), static_cast<const TT*>(warm_den));
    });
    return DFSPH_OK;
}

int32_t orc_prime(void* h, void* q, void* qd, const void* guess, double dt, double tol, int32_t max_it,
                  int32_t* iters) {
    if (!(dt > 0.0)) return fail(DFSPH_EINVAL, "dt must be > 0");
    DISPATCH(h, {
        using TT = std::remove_pointer_t<decltype(o->q)>;
        bind(o, q, qd, nullptr);
        *iters = o->prime(static_cast<const TT*>(guess), dt, tol, max_it);
    });
    return DFSPH_OK;
}

int32_t orc_build_neighbors(void* h, void* q) {
    DISPATCH(h, { bind(o, q, o->qd, nullptr); o->build_neighbors(); });
    return DFSPH_OK;
}

int32_t orc_boundary_slots(void* h, void* q, void* qd) {
    DISPATCH(h, { bind(o, q, qd, nullptr); o->boundary_slots(); });
    return DFSPH_OK;
}

int32_t orc_density_alpha(void* h, void* q) {
    DISPATCH(h, { bind(o, q, o->qd, nullptr); o->density_alpha(); });
    return DFSPH_OK;
}

int32_t orc_solve(void* h, int32_t mode, void* q, void* qd, double dt, int32_t fixed_iters) {
    if (!(dt > 0.0)) return fail(DFSPH_EINVAL, "dt must be > 0");
    DISPATCH(h, { bind(o, q, qd, nullptr); o->solve(mode, dt, fixed_iters); });
    return DFSPH_OK;
}

int32_t orc_flow_speed(void* h, void* qd) {
    DISPATCH(h, { bind(o, o->q, qd, nullptr); o->flow_speed(); });
    return DFSPH_OK;
}

int32_t orc_nonpressure(void* h, void* q, void* qd, const void* f, double dt) {
    DISPATCH(h, {
        using TT = std::remove_pointer_t<decltype(o->q)>;
        bind(o, q, qd, f);
        o->nonpressure(TT(dt));
        if (o->visc_on) o->visc_solve(TT(dt), true);
    });
    return DFSPH_OK;
}

int32_t orc_advect_project(void* h, void* q, void* qd, double dt) {
    DISPATCH(h, {
        using TT = std::remove_pointer_t<decltype(o->q)>;
        bind(o, q, qd, nullptr);
        o->advect_project(TT(dt));
    });
    return DFSPH_OK;
}

int32_t orc_meters(void* h) {
    DISPATCH(h, { o->meters(); });
    return DFSPH_OK;
}

int32_t orc_set_emitters(void* h, const DfsphEmitters* t) {
    if (!t) return fail(DFSPH_EINVAL, "null argument");
    DISPATCH(h, { return o->set_emitters(t); });
    return DFSPH_OK;
}

int32_t orc_set_emitter_enabled(void* h, int32_t e, int32_t on) {
    DISPATCH(h, {
        if (e < 0 || e >= o->E) return fail(DFSPH_EINVAL, "set_emitter_enabled: emitter does not exist");
        o->enabled[e] = on != 0;
    });
    return DFSPH_OK;
}

int32_t orc_emitters_step(void* h, void* q, void* qd, double dt) {
    DISPATCH(h, {
        using TT = std::remove_pointer_t<decltype(o->q)>;
        bind(o, q, qd, nullptr);
        o->emitters_step(TT(dt));
    });
    return DFSPH_OK;
}

int32_t orc_active_mass(void* h) {
    DISPATCH(h, { o->active_mass(); });
    return DFSPH_OK;
}

int32_t orc_zero_impulse(void* h) {
    DISPATCH(h, { std::fill(o->s_imp.begin(), o->s_imp.end(), 0); });
    return DFSPH_OK;
}

int32_t orc_get_buffer(void* h, int32_t which, void** ptr, int64_t* nb) {
    if (!ptr || !nb) return fail(DFSPH_EINVAL, "null argument");
    DISPATCH(h, { return get_buffer_t(o, which, ptr, nb); });
    return DFSPH_OK;
}

// host reductions for the reducer tests: values [n * ncomp] (T), out [W * ncomp] f64, mx [W] (T) or NULL
int32_t orc_reduce(void* h, const void* values, int32_t ncomp, double* out, void* mx) {
    DISPATCH(h, {
        using TT = std::remove_pointer_t<decltype(o->q)>;
        const TT* v = static_cast<const TT*>(values);
        std::vector<TT> vals(v, v + size_t(std::max(o->n, 1)) * ncomp);
        std::vector<double> res(size_t(o->W) * ncomp);
        o->reduce(vals, ncomp, res, static_cast<TT*>(mx));
        std::copy(res.begin(), res.end(), out);
    });
    return DFSPH_OK;
}

// scalar closed forms for the kernel tests (float or double per precision)
double orc_kernel_w(int32_t precision, double r, double h, int32_t kind) {
    return precision == 64 ? kernel_w<double>(r, h, kind) : double(kernel_w<float>(float(r), float(h), kind));
}

void orc_kernel_grad(int32_t precision, const double* rv, double h, int32_t kind, double* out) {
    if (precision == 64) {
        V3<double> g = kernel_grad<double>({rv[0], rv[1], rv[2]}, h, kind);
        out[0] = g.x, out[1] = g.y, out[2] = g.z;
    } else {
        V3<float> g = kernel_grad<float>({float(rv[0]), float(rv[1]), float(rv[2])}, float(h), kind);
        out[0] = g.x, out[1] = g.y, out[2] = g.z;
    }
}

void orc_halfspace_volume(int32_t precision, double d, double h, double H, int32_t kind, double* out) {
    if (precision == 64) {
        halfspace_volume<double>(d, h, H, kind, out[0], out[1]);
    } else {
        float a, b;
        halfspace_volume<float>(float(d), float(h), float(H), kind, a, b);
        out[0] = a, out[1] = b;
    }
}

void orc_shape_sdf(int32_t geo, const double* scale, const double* p, double* out) {
    double phi;
    V3<double> g;
    shape_sdf<double>(geo, {scale[0], scale[1], scale[2]}, {p[0], p[1], p[2]}, phi, g);
    out[0] = phi, out[1] = g.x, out[2] = g.y, out[3] = g.z;
}

// boundary_slots.py:50-102 on one float32 slot row: insert (s, phi, world normal nrm) into the phi-ascending row of
// n <= 4 slots (shape[4], phi[4], nrm[4 * 3], updated in place); returns 1 if the new wall was kept
int32_t orc_shape_insert(int32_t* shape, float* phi, float* nrm, int32_t n, int32_t s, float p, const float* n_new,
                         float near_) {
    Oracle<float> o;
    o.s_shape.assign(shape, shape + 4);
    o.s_phi.assign(phi, phi + 4);
    o.s_nbl.assign(nrm, nrm + 12);
    int kept = o.shape_insert(0, n, s, p, V3<float>{n_new[0], n_new[1], n_new[2]}, near_);
    std::copy(o.s_shape.begin(), o.s_shape.end(), shape);
    std::copy(o.s_phi.begin(), o.s_phi.end(), phi);
    std::copy(o.s_nbl.begin(), o.s_nbl.end(), nrm);
    return kept;
}

double orc_colour_distance(void* h, double c) {
    double r = 0.0;
    if (static_cast<Handle*>(h)->precision == 64) r = impl<double>(h)->colour_distance(c);
    else r = impl<float>(h)->colour_distance(float(c));
    return r;
}

}  // extern "C"
