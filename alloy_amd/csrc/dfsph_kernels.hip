// gfx950 kernels of the DFSPH liquid step (one HIP stream, deterministic, no float atomics).
//
// Layout in HBM. Caller-owned state: q, qd as [n][3] float32 (Newton's particle_q / particle_qd). Solver-owned
// work buffers, allocated once at create: per-particle float32 SoA fields; neighbour rows as column-major uint16
// world-local offsets nbr[k * n + i] (+ uint8 high byte for worlds over 65,536 particles), so the 64 lanes of a
// wavefront read 64 adjacent entries; boundary slots as [n][4] rows; per-world float64 reduction partials.
//
// Launch geometry. Per-particle phases that feed per-world reductions run one 256-thread block per "chunk": 256
// consecutive particles aligned to their world's start (reduce.py:167-182's chunk table). A block therefore
// belongs to one world: per-world control (converged, restart) is block-uniform, converged worlds' blocks exit at
// once (uneven batched worlds cost what their unconverged particles cost), and each block tree-reduces its values
// in float64 into a chunk partial. A per-world epilogue (one block per world) tree-reduces that world's chunk
// partials in fixed order and runs the world's control logic (check_kernel / cg_beta_kernel). The sums are thus
// bitwise independent of batching and of scheduling.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <type_traits>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <deque>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/dfsph_abi.h"
#include "dfsph_device.h"

using namespace dfsph;

void dfsph_launch_mlp_forward(const float* params, const float* X, int n, float* Y, const uint8_t* tile_on,
                              hipStream_t st);

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(DFSPH_EDEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// Everything a kernel reads: parameters as float32 kernel arguments and device pointers.
struct Ctx {
    int n, W, S, NW, cap, wide, kind, n_cells, friction, curvature, two_sided;
    int pair_cache;  // DFSPH_PAIR_CACHE_* bits: which gather passes read pgrad instead of recomputing the scale
    int sparse_corr; // DFSPH_OPT_SPARSE_CORRECTION: gated passes skip particles whose every term is an exact zero
    int n_chunks;
    float h, cell, r, n_full, tot_min, min_nb, omega, xsph, v_sep;
    float offset, amplitude, taper0, taper1, reach, eps, sweep_len;
    // model (caller-owned)
    const float* mass;
    const float* rho0;
    int* flags;
    int8_t* phase;
    const int* world;
    const float* gravity;
    const float* wvol;
    // tables
    const int* ws;
    const int* ch_begin;
    const int* ch_end;
    const int* ch_world;
    const int* w_chunk;
    int* ch_live;  // [n_chunks] 1 when the chunk holds liquid this step (k_groups_hist); the solve passes skip the
                   // other chunks, whose particles they would only write zeros for that no pass reads
    const float* g_org;
    const int* g_dims;
    const int* cell_base;
    const int* sh_start;
    const int* sh_ids;
    const int* sh_geo;
    const float* sh_scale;
    const float* sh_x;
    const float* sh_radius;
    const float* sh_mu;
    // grid
    int* cell_count;
    int* cell_start;
    uint8_t* wall_near;  // per cell, from wall_first: 1 when a wall particle is in it or one of its 26 neighbours
    int wall_first;
    int speed_later;     // the emitter pass (after emission) computes the speed statistics instead of the advect pass
    int one_way;         // DFSPHConfig.coupling == "one_way": the liquid does not push back on particle walls
    const int* wcidx;    // [n] compact index of each particle-wall particle (ascending ids), -1 otherwise
    long long* pw_exch;  // [NW*3] impulse of the exchange window per wall particle, 2^50 N s
    long long* pw_work;  // [W] the walls' work on the liquid in the last step, 2^50 J
    float* pw_col;       // [n] the particle-wall colour of each liquid particle, from the step's slot pass
    // the wall particles the slot pass's colour sum visited, in its order: [PW_ROW][n] compact wall index and
    // V_j W_ij, and the count (-1: more than PW_ROW; the reaction split then walks the cells itself)
    int* wr_id;
    float* wr_w;
    int* wr_n;
    // Verlet-skin reuse (neighbors_skin.py): superset rows within h + skin at each world's last build
    float skin;          // 0: off
    int cap_s;           // superset row capacity (<= 64)
    uint16_t* sup;       // [cap_s][n]
    uint8_t* sup_hi;     // [cap_s][n] when wide, else [1]
    int* sup_count;      // [n]; -1: this step's row was built plain (the superset overflowed)
    float4* q_ref;       // [n] position at the world's last build
    int* g_ref;          // [n] group at the world's last build (-2: never built)
    int* skin_flag;      // [W] the world rebuilds this step
    int* skin_force;     // [W] the world rebuilds next step
    // Weiler 2018 implicit viscosity (viscosity.py): per-world preconditioned CG on the velocities
    const float* mu;     // [n] sph:viscosity [Pa s]
    float visc_d;        // 2 (dim + 2) / the rest-lattice factor of the operator
    float *vb, *vx, *vr, *vz, *vp, *vap;  // [3n] right-hand side v*, solution, residual, preconditioned residual, ...
    float* vminv;        // [9n] inverse of each particle's 3x3 diagonal block
    int *v_conv, *v_iters;                // [W]
    double *v_rz, *v_bb, *v_step;         // [W]
    // its wall term (wall_viscosity.py), present when wall_coeff > 0
    float wall_coeff, mu_b;               // mu_b x wall_coefficient_scale; mu_b [Pa s]
    float* wall_m;                        // [9n] sum over the slots of the wall block M_ik
    float* wall_w;                        // [3n] sum over the slots of w_ik = M_ik v_b
    float *vshift, *vraw, *vgk;           // [3n], [3n], [n] the pressure guess's wall-velocity shift and its inputs
    // continuum surface force with a contact angle (csf.py, csf_sparse.py), present when csf != 0
    int csf;
    float csf_cs, csf_inv_i, csf_lo, csf_hi, csf_kcoh;  // c_s, 1 / I, CSF_BAND, 32 / (pi h^9)
    const float *sigma, *gamma;           // [n] surface tension [N/m], the pair force's gamma
    float *sh_cos, *sh_sin;               // [S] contact angle (sin -1: free)
    float *t_mcount, *t_kraw, *t_ktmp, *t_kappa;           // [n]
    float *t_grad, *t_g1, *t_g2, *t_nhat, *t_nwall, *t_anorm;  // [3n]
    int *t_surf, *t_wslot;                // [n]
    // LDS-staged halos of the gather passes (DFSPH_OPT_HALO): per chunk the sorted distinct neighbours of its
    // particles (global ids, at most HALO_MAX, h_count -1 when the chunk's halo does not fit) and per row entry the
    // neighbour's index in its chunk's halo ([cap + NB_U][n], the rows' layout and order)
    int use_halo;
    int* h_count;
    int* h_ids;
    uint16_t* h_slot;
    // packed gathers of the pressure acceleration: [3][n] {x, y, z, vec / rho0} for the three vectors k_accel
    // reads (kappa, the CG direction, the CG correction), and the per-world particle mass they stand in for
    int packed;
    float4* qk;
    float4* qa;  // packed (DFSPH_OPT_PACKED_ACCEL): {x, y, z, acc.x} and {acc.y, acc.z} of the last pressure
    float2* a2;  // acceleration, the operator apply's 24 B neighbour record instead of q4 and acc4 (32 B)
    const float* w_mass;
    int* arrive;  // [n] a grouped particle's arrival slot in its cell (the count's atomic returns it), for the scatter
    int* items;
    int* key;
    int* scan_part;
    // lists
    uint16_t* nbr;
    uint8_t* nbr_hi;
    float* pgrad;  // [cap][n] the step's gradient scale kernel_grad_scale(x_i - x_j) of every listed pair
    int* count;
    int* group;
    // slots
    int* s_shape;
    float* s_psi;
    float* s_gpsi;
    float* s_xb;
    float* s_vb;
    float* s_phi;
    float* s_xbl;
    float* s_nbl;
    float* s_imp;
    // per particle
    float *ratio, *alpha, *neff, *active, *reg, *src, *kappa, *err, *lower, *kv, *kd, *vtmp;
    float4* q4;    // (x, y, z, mass) per particle, refreshed from q at every neighbour build: one 16 B gather
    float4* q4s;   // (x, y, z, id bits) in cell order: the list build streams candidates from it
    int* items_u;     // the scatter's unordered cell occupants, then placed in ascending id order (k_place_cells)
    float4* q4s_u;
    float4* vrho;     // (v, ratio rho0) per particle: the XSPH pass's one 16 B record besides the position
    float4* acc4;  // pressure acceleration (x, y, z, 0)
    float* feat;     // [n][16] learned warm-start features (13 used)
    float* mlp_out;  // [n][2] network outputs
    float* warm_k;   // [2][n] learned initial kappa of the divergence / density solve
    int* learn_on;   // [W] the step's per-world gate of the learned start (DFSPH_OPT_LEARNED_GATE)
    int learn_gate = 0;  // 1: the network's features, outputs and start only for worlds with learn_on
    float kscale;    // |g| h: the kappa scale of the features and targets [m^2/s^2]
    float warm_v, warm_d;  // the reference warm-start factors of the divergence / density solve
    float lat_diag;  // interior rest-lattice DFSPH diagonal [1/m^2]
    float *pd, *z, *wv, *ap, *corr;
    float *kr, *pdr, *corrr;  // kappa / rho0, pd / rho0, corr / rho0: the quotients pressure_accel gathers
    uint8_t* touched;         // gated accel pass: 1 where the acceleration was evaluated (not the exact-zero skip)
    uint8_t* act;  // [n] the projected CG's active set (1: the particle's kappa is free), per solve
    // per world
    int *converged, *iters, *restart;
    float *err_mean, *err_max, *tol_scale;
    double *n_active, *err_sum, *rz, *stepb, *sums, *flow_z, *flow_xy, *head;
    int *st_overflow, *st_projected, *st_swept, *st_nonfinite, *it_den, *it_div;
    float *st_max_speed, *st_den_err, *st_div_err, *st_den_err_max, *st_den_tol;
    unsigned long long* meter;
    double* part;
    int* n_unconv;
    // the device-side solve loop (k_cg_loop): its grid-barrier counter and a fault flag (a barrier wait that ran out)
    unsigned* loop_bar;
    int* loop_fault;
    unsigned long long* prof_bytes;  // [2][n_chunks][4] per-wave byte counts while profiling, else null
    // state (bound per call)
    float* q;
    float* qd;
    const float* f;
};

__device__ __forceinline__ bool is_fluid(const Ctx& c, int i) {
    return (c.flags[i] & ACTIVE) != 0 && int(c.phase[i]) >= 1;
}
__device__ __forceinline__ f3 Q(const Ctx& c, int i) { return ld3(c.q, i); }
__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ f3 V(const Ctx& c, int i) { return ld3(c.qd, i); }

// Row storage in quads: entry k of particle i sits at ((k / 4) n + i) 4 + k % 4, so the four entries of a
// neighbour batch are one 8 B load per lane (4 B for the high bytes), coalesced across the wave, instead of four 2 B
// loads. dfsph_get_buffer presents the rows in the reference's layout [k][n] (rows_view).
__device__ __forceinline__ int64_t row_at(const Ctx& c, int i, int k) {
    return ((int64_t(k >> 2) * c.n + i) << 2) + (k & 3);
}
template <bool WIDE>
__device__ __forceinline__ int nbr_at(const Ctx& c, int base, int i, int k) {
    const int64_t r = row_at(c, i, k);
    int j = base + int(c.nbr[r]);
    if (WIDE) j += int(c.nbr_hi[r]) << 16;
    return j;
}

// Neighbour loops with staged loads: NB_U row entries, then the NB_U neighbour records, then the arithmetic in
// ascending k, exactly the sequential order. The loads of a batch are in flight together instead of one dependent
// index -> record round trip per neighbour; entries past the lane's count are read too (stale world-local ids of the
// same world, or the zero padding rows) and their bodies are skipped.
#ifndef DFSPH_NB_U
#define DFSPH_NB_U 4  // neighbour records in flight per lane (build-time tuning; 4 measured fastest of 1..16)
#endif
constexpr int NB_U = DFSPH_NB_U;
constexpr int ROW_PAD = 16;  // zero rows past the row capacity of nbr, nbr_hi, pgrad and h_slot
static_assert(NB_U % 4 == 0 && ROW_PAD >= 2 * NB_U, "a neighbour loop loads the next batch's row quads ahead");
template <class T>
__device__ __forceinline__ float4 load_pos(const T& d) { return d.q; }
[[maybe_unused]] __device__ __forceinline__ float4 load_pos(float4 d) { return d; }

// a vec / rho0 value of slot 0 (kappa), 1 (CG direction) or 2 (CG correction), and its packed copy
enum VrSlot { VR_KAPPA = 0, VR_DIR = 1, VR_CORR = 2 };
__device__ __forceinline__ void put_vr(const Ctx& c, float* vr, int slot, int i, float v) {
    vr[i] = v;
    if (c.packed) c.qk[int64_t(slot) * c.n + i].w = v;
}
// the filled boundary slots of particle i as a 4-bit mask (bit b: slot b holds a wall), for a pass to read before
// its neighbour loop: the loop's loads are ordered before what follows them, so a load after the loop would wait
__device__ __forceinline__ unsigned slot_mask(const Ctx& c, int i) {
    const int4 sh = *reinterpret_cast<const int4*>(c.s_shape + 4 * int64_t(i));
    return (sh.x >= 0 ? 1u : 0u) | (sh.y >= 0 ? 2u : 0u) | (sh.z >= 0 ? 4u : 0u) | (sh.w >= 0 ? 8u : 0u);
}
// the pressure acceleration of particle i: acc4, or with the packed records (c.packed) their acc components only
// (dfsph_get_buffer unpacks acc4 on demand)
__device__ __forceinline__ void put_acc(const Ctx& c, int i, float ax, float ay, float az) {
    if (c.packed) {
        c.qa[i].w = ax;
        c.a2[i] = make_float2(ay, az);
    } else {
        c.acc4[i] = make_float4(ax, ay, az, 0.0f);
    }
}
__device__ __forceinline__ f3 acc_of(const Ctx& c, int i) {
    if (c.packed) {
        const float2 a = c.a2[i];
        return f3{c.qa[i].w, a.x, a.y};
    }
    return xyz(c.acc4[i]);
}
// The row arrays hold ROW_PAD >= NB_U rows beyond `cap` (zero, never written), so a batch starting below the count
// loads all of its NB_U entries unconditionally: the loads issue back to back instead of one branch-guarded round
// trip each.
static_assert(NB_U % 4 == 0, "a neighbour batch is whole row quads");
// the NB_U world-local entries k0.. (k0 a multiple of 4) of particle i's row
template <bool WIDE>
__device__ __forceinline__ void row_entries(const Ctx& c, int i, int k0, unsigned (&e)[NB_U]) {
#pragma unroll
    for (int q = 0; q < NB_U / 4; ++q) {
        const int64_t r = int64_t((k0 >> 2) + q) * c.n + i;
        const uint2 v = reinterpret_cast<const uint2*>(c.nbr)[r];
        e[4 * q] = v.x & 0xffffu;
        e[4 * q + 1] = v.x >> 16;
        e[4 * q + 2] = v.y & 0xffffu;
        e[4 * q + 3] = v.y >> 16;
        if (WIDE) {
            const unsigned hb = reinterpret_cast<const unsigned*>(c.nbr_hi)[r];
#pragma unroll
            for (int u = 0; u < 4; ++u) e[4 * q + u] |= ((hb >> (8 * u)) & 0xffu) << 16;
        }
    }
}
template <bool WIDE>
__device__ __forceinline__ void nbr_batch(const Ctx& c, int i, int cnt, int base, int k0, int (&j)[NB_U]) {
    unsigned e[NB_U];
    row_entries<WIDE>(c, i, k0, e);
#pragma unroll
    for (int u = 0; u < NB_U; ++u) j[u] = k0 + u < cnt ? base + int(e[u]) : i;
}
// The neighbour loops of the per-particle passes (density, source terms, XSPH, the solves' acceleration and operator
// apply), with the addressing taken off the vector ALU, which the solve passes saturate:
// - the row entries (and cached scales) come through buffer loads whose row offset is a scalar and whose lane offset
//   (i times the entry size) is computed once, instead of a 64-bit address per entry;
// - the records are gathered through buffer loads on a per-world resource (the caller's load(e), see WorldRecs) with
//   the world-local entry times the record size as the lane offset: one shift per record; the caller's block holds
//   one chunk, so the world and the resource are uniform; a lane past its count gathers the record its stale entry
//   names (the rows hold world-local ids of the same world, zero when never written, and the resource's bound returns
//   zeros past the array), and its body is skipped;
// - for_rows_g's body receives rv * s without the zero-vector select of pair_grad: for a cut-off pair (s = 0) each
//   product is a signed zero, and the bodies only add such terms to (or subtract them from) sums that start at +0,
//   which leaves every sum unchanged (x + -0 = x, +0 - +0 = +0, and a sum that starts at +0 never becomes -0 in round
//   to nearest); a body that needs the exact zero vector forms pair_grad itself.
// Bodies run in ascending k, the reference's sequential order. Offsets stay below 2^32 for n < 2^28 (dfsph_create
// checks).
constexpr int BUF_WORD3 = 0x00020000;  // gfx9 buffer resource word 3: untyped access, 32-bit data format
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), short(0), int(bytes), BUF_WORD3);
}
// a world's records of one per-particle array, addressed by world-local id; `base` (the world's first particle) is
// uniform over the wave, which readfirstlane tells the compiler, so the resource lives in scalar registers
struct WorldRecs {
    __amdgpu_buffer_rsrc_t r;
    template <class T>
    __device__ __forceinline__ static WorldRecs of(const T* arr, int base, int n) {
        const int b = __builtin_amdgcn_readfirstlane(base);
        return WorldRecs{buf_rsrc(arr + b, uint32_t(n - b) * uint32_t(sizeof(T)))};
    }
    __device__ __forceinline__ float4 f4(unsigned e) const {
        return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, e * 16u, 0, 0));
    }
    __device__ __forceinline__ float2 f2(unsigned e) const {
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, e * 8u, 0, 0));
    }
    __device__ __forceinline__ float f1(unsigned e) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, e * 4u, 0, 0));
    }
    // a 12 B record (x, y, z), as 8 + 4 B (the compiler merges them into one 12 B load; its 96-bit load builtin
    // returned the first word for the third in a probe, so it is not used)
    __device__ __forceinline__ f3 v3(unsigned e) const {
        const uint2 a = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, e * 12u, 0, 0));
        const unsigned b = __builtin_amdgcn_raw_buffer_load_b32(r, e * 12u + 8u, 0, 0);
        return f3{__builtin_bit_cast(float, a.x), __builtin_bit_cast(float, a.y), __builtin_bit_cast(float, b)};
    }
};
// The NB_U world-local entries k0.. (a multiple of 4) of particle i's row: one buffer load per row quad (plus one for
// the high bytes), row offsets scalar, the lane offset 8 i. row_raw issues the loads, row_unpack takes the entries
// out of the loaded words, and row_batch does both.
struct RowRaw {
    uint2 v[NB_U / 4];
    unsigned hb[NB_U / 4];
};
template <bool WIDE>
__device__ __forceinline__ RowRaw row_raw(const Ctx& c, int i, int k0) {
    const uint32_t n = uint32_t(c.n);
    const __amdgpu_buffer_rsrc_t r16 = buf_rsrc(c.nbr + int64_t(k0) * n, 2u * NB_U * n);
    const __amdgpu_buffer_rsrc_t r8 = buf_rsrc(c.nbr_hi + (WIDE ? int64_t(k0) * n : 0), NB_U * n);
    RowRaw w;
#pragma unroll
    for (int q = 0; q < NB_U / 4; ++q) {
        w.v[q] = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r16, 8u * unsigned(i),
                                                                                 int(8u * q * n), 0));
        w.hb[q] = WIDE ? __builtin_amdgcn_raw_buffer_load_b32(r8, 4u * unsigned(i), int(4u * q * n), 0) : 0u;
    }
    return w;
}
// the next batch's quads (k0 + NB_U..) where `next`, else batch k0's again: a lane that has no next batch re-reads
// lines it has just read instead of fetching a row past its count (within the ROW_PAD rows past the capacity)
template <bool WIDE>
__device__ __forceinline__ RowRaw row_raw_next(const Ctx& c, int i, int k0, bool next) {
    const uint32_t n = uint32_t(c.n);
    const __amdgpu_buffer_rsrc_t r16 = buf_rsrc(c.nbr + int64_t(k0) * n, 4u * NB_U * n);
    const __amdgpu_buffer_rsrc_t r8 = buf_rsrc(c.nbr_hi + (WIDE ? int64_t(k0) * n : 0), 2u * NB_U * n);
    const unsigned step = next ? unsigned(NB_U / 4) * n : 0u;  // quads of the next batch
    RowRaw w;
#pragma unroll
    for (int q = 0; q < NB_U / 4; ++q) {
        w.v[q] = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r16, 8u * (unsigned(i) + step),
                                                                                 int(8u * q * n), 0));
        w.hb[q] = WIDE ? __builtin_amdgcn_raw_buffer_load_b32(r8, 4u * (unsigned(i) + step), int(4u * q * n), 0)
                       : 0u;
    }
    return w;
}
template <bool WIDE>
__device__ __forceinline__ void row_unpack(const RowRaw& w, unsigned (&e)[NB_U]) {
#pragma unroll
    for (int q = 0; q < NB_U / 4; ++q) {
        e[4 * q] = w.v[q].x & 0xffffu;
        e[4 * q + 1] = w.v[q].x >> 16;
        e[4 * q + 2] = w.v[q].y & 0xffffu;
        e[4 * q + 3] = w.v[q].y >> 16;
        if (WIDE) {
#pragma unroll
            for (int u = 0; u < 4; ++u) e[4 * q + u] |= ((w.hb[q] >> (8 * u)) & 0xffu) << 16;
        }
    }
}
template <bool WIDE>
__device__ __forceinline__ void row_batch(const Ctx& c, int i, int k0, unsigned (&e)[NB_U]) {
    row_unpack<WIDE>(row_raw<WIDE>(c, i, k0), e);
}
// body(k, record) in ascending k over the world's records load(e). The row entries of batch k0 + NB_U are loaded
// while batch k0 computes (at the last batch from the ROW_PAD rows past the capacity), and every load of a batch is
// issued before its per-pair branches (the compiler would otherwise sink a load into the one branch that uses it,
// behind the previous pair's arithmetic).
template <bool WIDE, class T, class LoadF, class BodyF>
__device__ __forceinline__ void for_rows(const Ctx& c, int i, int cnt, LoadF load, BodyF body) {
    RowRaw raw = row_raw<WIDE>(c, i, 0);
    for (int k0 = 0; k0 < cnt; k0 += NB_U) {
        unsigned e[NB_U];
        row_unpack<WIDE>(raw, e);
        T d[NB_U];
#pragma unroll
        for (int u = 0; u < NB_U; ++u) d[u] = load(e[u]);
        raw = row_raw_next<WIDE>(c, i, k0, k0 + NB_U < cnt);
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int u = 0; u < NB_U; ++u)
            if (k0 + u < cnt) body(k0 + u, d[u]);
    }
}
// body(record, grad W_ij) in ascending k, the gradient from the pair cache (cached) or evaluated. The gradient's
// source is chosen once per loop (GS), not per pair: each version of the loop is then straight-line code over its
// NB_U pairs, which the compiler can interleave, where a per-pair launch-uniform branch would split it into blocks.
enum GradSource { GS_CACHED = 0, GS_CUBIC, GS_ANY };
template <int GS>
__device__ __forceinline__ float grad_scale_of(f3 rv, const KC& K) {
    if (GS == GS_CUBIC) return kernel_grad_scale_fast_kind<0>(rv, K);
    return kernel_grad_scale(rv, K);
}
template <bool WIDE, int GS, class T, class LoadF, class BodyF>
__device__ __forceinline__ void rows_g_loop(const Ctx& c, int i, f3 xi, int cnt, const KC& K, LoadF& load,
                                            BodyF& body) {
    const uint32_t n = uint32_t(c.n);
    RowRaw raw = row_raw<WIDE>(c, i, 0);  // pipelined as in for_rows
    for (int k0 = 0; k0 < cnt; k0 += NB_U) {
        unsigned e[NB_U];
        row_unpack<WIDE>(raw, e);
        float sc[NB_U];
        if (GS == GS_CACHED) {
            const __amdgpu_buffer_rsrc_t rg = buf_rsrc(c.pgrad + int64_t(k0) * n, 4u * NB_U * n);
#pragma unroll
            for (int u = 0; u < NB_U; ++u)
                sc[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rg, 4u * unsigned(i),
                                                                                       int(4u * u * n), 0));
        }
        T d[NB_U];
#pragma unroll
        for (int u = 0; u < NB_U; ++u) d[u] = load(e[u]);
        raw = row_raw_next<WIDE>(c, i, k0, k0 + NB_U < cnt);
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int u = 0; u < NB_U; ++u) {
            if (k0 + u >= cnt) continue;
            const f3 rv = xi - xyz(load_pos(d[u]));
            body(d[u], rv * (GS == GS_CACHED ? sc[u] : grad_scale_of<GS>(rv, K)));
        }
    }
}
template <bool WIDE, class T, class LoadF, class BodyF>
__device__ __forceinline__ void for_rows_g(const Ctx& c, int i, f3 xi, int cnt, bool cached, LoadF load,
                                           BodyF body) {
    const KC K = make_kc(c.h, c.kind);
    if (cached) rows_g_loop<WIDE, GS_CACHED, T>(c, i, xi, cnt, K, load, body);
    else if (K.fast && K.kind == 0) rows_g_loop<WIDE, GS_CUBIC, T>(c, i, xi, cnt, K, load, body);
    else rows_g_loop<WIDE, GS_ANY, T>(c, i, xi, cnt, K, load, body);
}

#define LAUNCH_W(kernel, grid, block, ...)                                        \
    do {                                                                          \
        if (c.wide) kernel<true><<<(grid), (block), 0, st>>>(__VA_ARGS__);        \
        else kernel<false><<<(grid), (block), 0, st>>>(__VA_ARGS__);             \
    } while (0)

// kernel ids of the per-launch profiling (dfsph_set_profiling / dfsph_profile_read)
enum KernelId { KID_ACCEL = 0, KID_APPLY, KID_BUILD, KID_DENSITY, KID_SLOTS, KID_NONPRESSURE, KID_ADVECT,
                KID_SOLVE_INIT, KID_POST, KID_COUNT };

// Profiling only (c.prof_bytes set): add the algorithmic bytes of this wave's particles to the wave's own slot of
// kernel kid's counters ([2][4 n_chunks]: k_accel, k_apply), summed on the host. One plain add per wave and no
// shared address: a single counter would take one atomic per wave from every block and serialise the launch.
// Called where every lane of a block that did work is still active.
__device__ __forceinline__ void account_bytes(const Ctx& c, int kid, int chunk, unsigned v) {
    unsigned long long s = v;
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if ((threadIdx.x & 63) == 0) {
        unsigned long long* slot = c.prof_bytes + (int64_t(kid) * c.n_chunks + chunk) * 4 + (threadIdx.x >> 6);
        *slot += s;
    }
}

// ----------------------------------------------------------------------------------------------- XCD mapping
// The dispatcher hands workgroup b to XCD b % 8. Renumbering so that XCD x runs the contiguous logical range
// [x q + min(x, r), ...) (G = 8 q + r) keeps the particles whose neighbours a block gathers inside one XCD's L2
// instead of spreading consecutive chunks over all eight. A bijection of [0, G): results do not depend on it.
constexpr int N_XCD = 8;
__device__ __forceinline__ int xcd_block(int b, int G) {
    int q = G / N_XCD, r = G % N_XCD, x = b % N_XCD;
    return x * q + min(x, r) + b / N_XCD;
}
__device__ __forceinline__ int xcd_particle() {
    return xcd_block(int(blockIdx.x), int(gridDim.x)) * int(blockDim.x) + int(threadIdx.x);
}

// ----------------------------------------------------------------------------------------------- reductions
// Fixed 256-wide tree: a[t] += a[t + s] for s = 128, 64 (LDS) then 32 .. 1 (wavefront shuffles; the same pairs).
__device__ __forceinline__ double tree_sum(double v, double* lds) {
    int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    if (t < 128) lds[t] += lds[t + 128];
    __syncthreads();
    if (t < 64) lds[t] += lds[t + 64];
    __syncthreads();
    double r = 0.0;
    if (t < 64) {
        r = lds[t];
        for (int s = 32; s >= 1; s >>= 1) r += __shfl_down(r, s, 64);
    }
    __syncthreads();
    return r;  // valid in thread 0
}
__device__ __forceinline__ double tree_max(double v, double* lds) {
    int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    if (t < 128) lds[t] = fmax(lds[t], lds[t + 128]);
    __syncthreads();
    if (t < 64) lds[t] = fmax(lds[t], lds[t + 64]);
    __syncthreads();
    double r = -3.4e38;
    if (t < 64) {
        r = lds[t];
        for (int s = 32; s >= 1; s >>= 1) r = fmax(r, __shfl_down(r, s, 64));
    }
    __syncthreads();
    return r;
}
// chunk b_'s partial (sum x, sum y, max y) of the block's values; invalid lanes pass (0, 0, -3.4e38)
__device__ __forceinline__ void chunk_reduce(const Ctx& c, int b_, double x, double y, double m) {
    __shared__ double lds[CHUNK];
    double sx = tree_sum(x, lds);
    double sy = tree_sum(y, lds);
    double mx = tree_max(m, lds);
    if (threadIdx.x == 0) {
        double* p = c.part + 3 * int64_t(b_);
        p[0] = sx;
        p[1] = sy;
        p[2] = mx;
    }
}

// chunk_reduce(c, b_, x, 0.0, valid ? 0.0 : -3.4e38) with one tree: the second sum of +0 terms is +0, and the max is
// 0.0 when a lane of the block is valid, else -3.4e38
__device__ __forceinline__ void chunk_reduce_x(const Ctx& c, int b_, double x, bool valid) {
    __shared__ double lds[CHUNK];
    const double sx = tree_sum(x, lds);
    const bool any = __syncthreads_or(valid ? 1 : 0) != 0;
    if (threadIdx.x == 0) {
        double* p = c.part + 3 * int64_t(b_);
        p[0] = sx;
        p[1] = 0.0;
        p[2] = any ? 0.0 : -3.4e38;
    }
}

// The chunk partial of a chunk without liquid: every term is +0 and every max term 0.0 (each chunk has a particle)
__device__ __forceinline__ void chunk_zero(const Ctx& c, int b_) {
    if (threadIdx.x == 0) {
        double* p = c.part + 3 * int64_t(b_);
        p[0] = 0.0;
        p[1] = 0.0;
        p[2] = 0.0;
    }
}

enum WorldMode { W_NACTIVE = 0, W_CG_INIT, W_CG_PAP, W_CG_BETA, W_JACOBI, W_FLOW_Z, W_FLOW_XY, W_HEAD, W_VISC_BEGIN,
                 W_VISC_ALPHA, W_VISC_CHECK };

struct StopSpec {
    float tol, tol_max;
    int min_it, max_it, use_scale;
};

__device__ __forceinline__ bool stop_test(const Ctx& c, int w, int it, float mean, const StopSpec& sp) {
    float ts = sp.use_scale ? c.tol_scale[w] : 1.0f;
    return (it >= sp.min_it && mean <= sp.tol * ts && c.err_max[w] <= sp.tol_max * ts) || it >= sp.max_it;
}

// One block per world: tree over the world's chunk partials (thread t sums chunks t, t + 256, ... in order), then
// the world's control logic in thread 0 (reduce.py:100-164 + pressure.py:370-390 / krylov.py:215-240).
__device__ void world_body(const Ctx& c, int w, int mode, const StopSpec& sp) {
    if (mode == W_CG_PAP && w == 0 && threadIdx.x == 0) *c.n_unconv = 0;  // recounted by W_CG_BETA
    if ((mode == W_CG_PAP || mode == W_CG_BETA || mode == W_JACOBI) && c.converged[w] != 0) return;
    if (mode == W_VISC_ALPHA && w == 0 && threadIdx.x == 0) *c.n_unconv = 0;  // recounted by W_VISC_CHECK
    if ((mode == W_VISC_ALPHA || mode == W_VISC_CHECK) && c.v_conv[w] != 0) return;
    __shared__ double lds[CHUNK];
    double a = 0.0, b = 0.0, m = -3.4e38;
    for (int k = c.w_chunk[w] + int(threadIdx.x); k < c.w_chunk[w + 1]; k += CHUNK) {
        a += c.part[3 * int64_t(k)];
        b += c.part[3 * int64_t(k) + 1];
        m = fmax(m, c.part[3 * int64_t(k) + 2]);
    }
    // the second sum and the max only for the modes that read them (the chunk partials of the others carry +0 there)
    const bool need_b = !(mode == W_NACTIVE || mode == W_CG_INIT || mode == W_CG_PAP || mode == W_JACOBI ||
                          mode == W_VISC_ALPHA);
    const bool need_m = mode == W_CG_BETA || mode == W_JACOBI;
    double sa = tree_sum(a, lds);
    double sb = need_b ? tree_sum(b, lds) : 0.0;
    double mx = need_m ? tree_max(m, lds) : -3.4e38;
    if (threadIdx.x != 0) return;
    switch (mode) {
        case W_NACTIVE:
            c.n_active[w] = sa;
            break;
        case W_CG_INIT:
            c.rz[w] = sa;
            c.restart[w] = 0;
            break;
        case W_CG_PAP:
            c.sums[2 * w] = sa;
            c.sums[2 * w + 1] = sb;
            break;
        case W_CG_BETA: {
            c.err_max[w] = float(fmax(mx, 0.0));
            int it = c.iters[w] + 1;
            c.iters[w] = it;
            double rz_new = sa;
            float mean = float(sb / fmax(c.n_active[w], 1.0));
            c.err_mean[w] = mean;
            double beta = 0.0;
            if (c.restart[w] == 0 && c.rz[w] > 1.0e-30) beta = rz_new / c.rz[w];
            c.restart[w] = 0;
            c.rz[w] = rz_new;
            c.stepb[w] = beta;
            if (stop_test(c, w, it, mean, sp)) c.converged[w] = 1;
            else atomicAdd(c.n_unconv, 1);
            break;
        }
        case W_JACOBI: {
            c.err_sum[w] = sa;
            c.err_max[w] = float(fmax(mx, 0.0));
            int it = c.iters[w] + 1;
            c.iters[w] = it;
            float mean = float(sa / fmax(c.n_active[w], 1.0));
            c.err_mean[w] = mean;
            if (stop_test(c, w, it, mean, sp)) c.converged[w] = 1;
            else atomicAdd(c.n_unconv, 1);
            break;
        }
        case W_FLOW_Z:
            c.flow_z[2 * w] = sa;
            c.flow_z[2 * w + 1] = sb;
            break;
        case W_FLOW_XY:
            c.flow_xy[2 * w] = sa;
            c.flow_xy[2 * w + 1] = sb;
            break;
        case W_HEAD:
            c.head[2 * w] = sa;
            c.head[2 * w + 1] = sb;
            break;
        case W_VISC_BEGIN:  // viscosity.py:120-132
            c.v_rz[w] = sa;
            c.v_bb[w] = sb;
            c.v_iters[w] = 0;
            if (c.n_active[w] > 0.0 && sa > 0.0) {
                c.v_conv[w] = 0;
                atomicAdd(c.n_unconv, 1);
            } else {
                c.v_conv[w] = 1;
            }
            break;
        case W_VISC_ALPHA: {  // viscosity.py:135-145
            double a = 0.0;
            if (sa > 1.0e-30) a = c.v_rz[w] / sa;
            c.v_step[w] = a;
            break;
        }
        case W_VISC_CHECK: {  // viscosity.py:166-185: stop at |r|^2 <= tol^2 |b|^2; beta into v_step
            int it = c.v_iters[w] + 1;
            c.v_iters[w] = it;
            double beta = 0.0;
            if (c.v_rz[w] > 1.0e-30) beta = sa / c.v_rz[w];
            c.v_rz[w] = sa;
            c.v_step[w] = beta;
            float tol = sp.tol;
            if (sb <= double(tol * tol) * c.v_bb[w] || it >= sp.max_it) c.v_conv[w] = 1;
            else atomicAdd(c.n_unconv, 1);
            break;
        }
    }
}

__global__ void __launch_bounds__(256) k_world(Ctx c, int mode, StopSpec sp) { world_body(c, blockIdx.x, mode, sp); }

// ----------------------------------------------------------------------------------------------- utilities
__global__ void k_fill_i32(int* p, int v, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
__global__ void k_fill_f32(float* p, float v, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
__global__ void k_copy_f32(float* dst, const float* src, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}
// dst = a * src (prime.py: kv = guess * dt, kd = dt * kv)
__global__ void k_scale_f32(float* dst, const float* src, float a, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = a * src[i];
}
__global__ void k_scale_f32_rev(float* dst, const float* src, float a, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] * a;
}

// ----------------------------------------------------------------------------------------------- neighbour grid
__device__ __forceinline__ int cell_coord(float x, float o, float cell, int dim) {
    float u = (x - o) / cell;
    float fl = floorf(u);
    if (!(fl >= 0.0f)) return 0;
    if (fl >= float(dim - 1)) return dim - 1;
    return int(fl);
}
// cell of x in grid g (g < W: fluid grid of world g; g >= W: wall grid of world g - W)
__device__ __forceinline__ int cell_of(const Ctx& c, int g, f3 x, int& cx, int& cy, int& cz) {
    int w = g < c.W ? g : g - c.W;
    const int* d = c.g_dims + 3 * w;
    cx = cell_coord(x.x, c.g_org[3 * w], c.cell, d[0]);
    cy = cell_coord(x.y, c.g_org[3 * w + 1], c.cell, d[1]);
    cz = cell_coord(x.z, c.g_org[3 * w + 2], c.cell, d[2]);
    return c.cell_base[g] + (cx * d[1] + cy) * d[2] + cz;  // z fastest: z-rows of cells are contiguous
}

// neighbors.py:36-51 group per particle, then its cell key and the cell histogram
__global__ void __launch_bounds__(256) k_groups_hist(Ctx c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int k = -1;
    if (i < c.n) {
        int g = -1;
        int fl = c.flags[i];
        if ((fl & ACTIVE) != 0 && int(c.phase[i]) >= 1) g = c.world[i];
        else if (c.NW > 0 && c.wvol[i] > 0.0f && (fl & ACTIVE) != 0 && c.world[i] >= 0) g = -2 - c.world[i];
        c.group[i] = g;
        if (g >= 0) c.ch_live[c.w_chunk[g] + (i - c.ws[g]) / CHUNK] = 1;
        f3 xi = Q(c, i);
        c.q4[i] = make_float4(xi.x, xi.y, xi.z, c.mass[i]);
        if (c.packed && g >= 0) {  // liquid only (the rows list liquid); w is written by the producer before any read
            for (int s = 0; s < 3; ++s) c.qk[int64_t(s) * c.n + i] = make_float4(xi.x, xi.y, xi.z, 0.0f);
            c.qa[i] = make_float4(xi.x, xi.y, xi.z, 0.0f);
        }
        if (g != -1) {
            int gi = g >= 0 ? g : c.W + (-2 - g);
            int cx, cy, cz;
            k = cell_of(c, gi, xi, cx, cy, cz);
        }
        c.key[i] = k;
    }
    // arrival slots: consecutive lanes in one cell (neighbouring ids often share a cell) form a run whose first lane
    // takes the run's slots with one atomic; each lane's slot is the run's base plus its place in the run
    const int lane = int(threadIdx.x) & 63;
    const int prev = __shfl_up(k, 1, 64);
    const bool head = lane == 0 || prev != k;
    const unsigned long long heads = __ballot(head);
    const int first = 63 - __clzll(static_cast<long long>(heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull))));
    const unsigned long long after = first == 63 ? 0ull : heads & ~((2ull << first) - 1ull);
    const int end = after ? __ffsll(static_cast<long long>(after)) - 1 : 64;
    int base = 0;
    if (head && k >= 0) base = atomicAdd(c.cell_count + k, end - first);
    base = __shfl(base, first, 64);
    if (k >= 0) c.arrive[i] = base + (lane - first);
}

// exclusive scan of a block's 1024 elements (256 threads x 4); returns the thread's exclusive prefix, total in *tot
__device__ __forceinline__ int block_scan_excl(int v, int* tot) {
    __shared__ int wsum[4];
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int d = 1; d < 64; d <<= 1) {
        int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    int off = 0;
    for (int k = 0; k < wid; ++k) off += wsum[k];
    *tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return off + x - v;
}

// The cell scan in three passes over blocks of SCAN_B cells (SCAN_T consecutive cells per thread, loaded as int4
// where the block lies inside the array): block totals, an exclusive scan of the totals in one block, then each
// block's exclusive prefix.
constexpr int SCAN_T = 16, SCAN_B = 256 * SCAN_T;
__device__ __forceinline__ void scan_load(const int* in, int n, int base, int (&v)[SCAN_T]) {
    if (base + SCAN_T <= n) {
        const int4* p = reinterpret_cast<const int4*>(in + base);
#pragma unroll
        for (int q = 0; q < SCAN_T / 4; ++q) {
            const int4 x = p[q];
            v[4 * q] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < SCAN_T; ++k) v[k] = base + k < n ? in[base + k] : 0;
    }
}
__global__ void __launch_bounds__(256) k_scan_reduce(const int* in, int n, int* part) {
    const int base = blockIdx.x * SCAN_B + threadIdx.x * SCAN_T;
    int v[SCAN_T], s = 0;
    scan_load(in, n, base, v);
#pragma unroll
    for (int k = 0; k < SCAN_T; ++k) s += v[k];
    int tot;
    block_scan_excl(s, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}
// single block: exclusive scan of the block totals in place; total into out[n]
__global__ void __launch_bounds__(256) k_scan_top(int* part, int nb, int* out, int n) {
    int running = 0;
    for (int b0 = 0; b0 < nb; b0 += 1024) {
        int base = b0 + threadIdx.x * 4;
        int v[4], s = 0;
        for (int k = 0; k < 4; ++k) {
            v[k] = base + k < nb ? part[base + k] : 0;
            s += v[k];
        }
        int tot;
        int ex = block_scan_excl(s, &tot) + running;
        for (int k = 0; k < 4; ++k) {
            if (base + k < nb) part[base + k] = ex;
            ex += v[k];
        }
        running += tot;
    }
    if (threadIdx.x == 0) out[n] = running;
}
__global__ void __launch_bounds__(256) k_scan_apply(const int* in, int n, const int* part, int* out) {
    const int base = blockIdx.x * SCAN_B + threadIdx.x * SCAN_T;
    int v[SCAN_T], s = 0;
    scan_load(in, n, base, v);
#pragma unroll
    for (int k = 0; k < SCAN_T; ++k) s += v[k];
    int tot;
    int ex = block_scan_excl(s, &tot) + part[blockIdx.x];
    if (base + SCAN_T <= n) {
        int4* p = reinterpret_cast<int4*>(out + base);
#pragma unroll
        for (int q = 0; q < SCAN_T / 4; ++q) {
            int4 x;
            x.x = ex;
            ex += v[4 * q];
            x.y = ex;
            ex += v[4 * q + 1];
            x.z = ex;
            ex += v[4 * q + 2];
            x.w = ex;
            ex += v[4 * q + 3];
            p[q] = x;
        }
    } else {
        for (int k = 0; k < SCAN_T; ++k) {
            if (base + k < n) out[base + k] = ex;
            ex += v[k];
        }
    }
}

__global__ void __launch_bounds__(256) k_scatter(Ctx c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    int k = c.key[i];
    if (k < 0) return;
    const int pos = c.cell_start[k] + c.arrive[i];
    c.items_u[pos] = i;
    float4 p = c.q4[i];
    c.q4s_u[pos] = make_float4(p.x, p.y, p.z, __int_as_float(i));
}

// ascending ids within each cell: the fixed visit order of the colour sums (wall cells) and a nearly sorted
// arrival order for the list build (fluid cells). One thread per occupant of the scatter's unordered slots: its rank
// in its cell is the number of the cell's occupants with a lower id, and it moves its id and record there.
__global__ void __launch_bounds__(256) k_place_cells(Ctx c) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= c.cell_start[c.n_cells]) return;
    const int v = c.items_u[a];
    const int cc = c.key[v];
    const int b = c.cell_start[cc], e = c.cell_start[cc + 1];
    int rank = 0;
    for (int k = b; k < e; ++k) rank += c.items_u[k] < v ? 1 : 0;
    c.items[b + rank] = v;
    c.q4s[b + rank] = c.q4s_u[a];
}

// Every wall particle marks the cells of its wall grid within one cell of its own (the block pw_sample walks), so a
// point whose cell is unmarked has no wall particle in its walk and its colour sums are exactly zero.
__global__ void __launch_bounds__(256) k_mark_wall_cells(Ctx c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    int g = c.group[i];
    if (g > -2) return;  // not a wall particle
    int w = -2 - g;
    int cx, cy, cz;
    cell_of(c, c.W + w, xyz(c.q4[i]), cx, cy, cz);
    const int* d = c.g_dims + 3 * w;
    int cb = c.cell_base[c.W + w];
    for (int X = max(cx - 1, 0); X <= min(cx + 1, d[0] - 1); ++X)
        for (int Y = max(cy - 1, 0); Y <= min(cy + 1, d[1] - 1); ++Y)
            for (int Z = max(cz - 1, 0); Z <= min(cz + 1, d[2] - 1); ++Z)
                c.wall_near[cb + (X * d[1] + Y) * d[2] + Z - c.wall_first] = 1;
}
__global__ void k_fill_u8(uint8_t* p, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = 0;
}

// The rows in the reference's layout, entry k of particle i at k n + i (neighbors.py:198-230), from the stored quads
// (to_view) or back into them: the buffer interface's view of the rows.
__global__ void k_rows_view(Ctx c, uint16_t* v16, uint8_t* v8, int to_view) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(c.cap) * c.n) return;
    const int k = int(t / c.n), i = int(t % c.n);
    const int64_t r = row_at(c, i, k);
    if (to_view) {
        v16[t] = c.nbr[r];
        if (c.wide) v8[t] = c.nbr_hi[r];
    } else {
        c.nbr[r] = v16[t];
        if (c.wide) c.nbr_hi[r] = v8[t];
    }
}

// neighbors.py:160-195: rows of the lowest `cap` world-local ids within h (fp32 dot(d, d) < h^2), ascending;
// overflow counted per world. Each thread keeps its row in LDS (column per thread: conflict-free), as uint16
// world-local ids unless a world exceeds 65,536 particles (WIDE: 32-bit entries), so the narrow build holds
// twice the waves per CU.
#ifndef DFSPH_BUILD_U
#define DFSPH_BUILD_U 8  // candidate records in flight per lane in the list build (build-time tuning)
#endif
constexpr int BUILD_U = DFSPH_BUILD_U;
// BB: the block size (DFSPH_OPT_BUILD_BLOCK); the row of thread t is row[k * BB + t]. One row: the lowest `cap`
// world-local ids within sqrt(r2) of particle i (group g >= 0), ascending, in LDS; returns (count, candidates).
// Accepted candidates are appended unsorted (a branch-free test per candidate, one LDS store per acceptance) and the
// row is sorted once at the end; a thread whose entries could overflow ROW_BUF within the next batch sorts them and
// keeps the lowest `cap`, so the result is the lowest `cap` of all candidates for any count. The candidate batches load unclamped: q4s holds
// BUILD_U records past its end, and records past the z-row's range are rejected by the test.
#ifndef DFSPH_ROW_BUF
#define DFSPH_ROW_BUF 80  // row entries collected per thread before a sort (build-time tuning; cap + BUILD_U at least)
#endif
constexpr int ROW_BUF = DFSPH_ROW_BUF;
static_assert(64 + BUILD_U <= ROW_BUF, "a compacted row and one candidate batch must fit the row buffer");
template <class RowT, int BB>
__device__ __forceinline__ void sort_row(RowT* L, int nn) {
    // insertion sort, ascending: the candidates arrive nearly sorted (ids ascending within a cell, cells in x, y, z
    // order), so most entries stay where they are
    for (int a = 1; a < nn; ++a) {
        const RowT v = L[a * BB];
        int k = a - 1;
        RowT u = L[k * BB];
        while (u > v) {
            L[(k + 1) * BB] = u;
            if (--k < 0) break;
            u = L[k * BB];
        }
        L[(k + 1) * BB] = v;
    }
}
template <bool WIDE, int BB>
__device__ __forceinline__ void build_row(const Ctx& c, int i, int g, f3 xi, float r2, int cap,
                                          typename std::conditional<WIDE, int, uint16_t>::type* L, int& nn_out,
                                          int& total_out) {
    using RowT = typename std::conditional<WIDE, int, uint16_t>::type;
    const float h2 = r2;
    int base = c.ws[g];
    int cx, cy, cz;
    cell_of(c, g, xi, cx, cy, cz);
    const int* d = c.g_dims + 3 * g;
    int nn = 0, total = 0;
    // x outer, y, then the contiguous z-row of 3 cells, ids ascending within a cell: the nine rows r = 3 (dx + 1) +
    // (dy + 1) in order, the range of row r + 1 loaded before row r's candidates (an empty range off the grid)
    const int z0 = max(cz - 1, 0), z1 = min(cz + 1, d[2] - 1);
    auto row_range = [&](int r, int& a, int& b) {
        const int X = cx + r / 3 - 1, Y = cy + r % 3 - 1;
        a = b = 0;
        if (X < 0 || X >= d[0] || Y < 0 || Y >= d[1]) return;
        const int cb = c.cell_base[g] + (X * d[1] + Y) * d[2];
        a = c.cell_start[cb + z0];
        b = c.cell_start[cb + z1 + 1];
    };
    int n0, n1;
    row_range(0, n0, n1);
    for (int r = 0; r < 9; ++r) {
        const int e0 = n0, e1 = n1;
        if (r < 8) row_range(r + 1, n0, n1);
        for (int e = e0; e < e1; e += BUILD_U) {
            if (nn > ROW_BUF - BUILD_U) {  // room for a whole batch: keep the lowest cap (cap + BUILD_U <= ROW_BUF)
                sort_row<RowT, BB>(L, nn);
                nn = cap;
            }
            const float4* pe = c.q4s + e;
            float4 pc[BUILD_U];
#pragma unroll
            for (int u = 0; u < BUILD_U; ++u) pc[u] = pe[u];
#pragma unroll
            for (int u = 0; u < BUILD_U; ++u) {
                const int j = __float_as_int(pc[u].w);
                const f3 dd = xyz(pc[u]) - xi;
                const bool ok = (e + u < e1) & (j != i) & (dot(dd, dd) < h2);
                // written whether accepted or not (a rejected entry is overwritten by the next), no branch
                L[nn * BB] = RowT(j - base);
                total += ok ? 1 : 0;
                nn += ok ? 1 : 0;
            }
        }
    }
    sort_row<RowT, BB>(L, nn);
    nn_out = nn < cap ? nn : cap;
    total_out = total;
}

template <bool WIDE, int BB>
__device__ __forceinline__ void write_row(const Ctx& c, int i, int g, int nn, int total,
                                          const typename std::conditional<WIDE, int, uint16_t>::type* L) {
    for (int k0 = 0; k0 < nn; k0 += 4) {  // whole quads; entries past nn are 0 (a valid world-local id)
        unsigned v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = k0 + u < nn ? unsigned(L[(k0 + u) * BB]) : 0u;
        const int64_t r = int64_t(k0 >> 2) * c.n + i;
        reinterpret_cast<uint2*>(c.nbr)[r] = make_uint2((v[0] & 0xffffu) | (v[1] << 16), (v[2] & 0xffffu) | (v[3] << 16));
        if (WIDE)
            reinterpret_cast<unsigned*>(c.nbr_hi)[r] = (v[0] >> 16) | ((v[1] >> 16) << 8) | ((v[2] >> 16) << 16) |
                                                       ((v[3] >> 16) << 24);
    }
    if (total > c.cap) atomicAdd(c.st_overflow + g, 1);
    c.count[i] = nn;
}

// SUP false: the step's rows (neighbors.py:160-195). SUP true (neighbors_skin.py:57-106): the superset rows within
// h + skin of the worlds flagged this step, each particle's reference position and group stored; a superset that
// overflows its capacity is replaced by this step's plain row and forces the world's next rebuild.
// Threads take the particles in id order, so the row writes of a wave are coalesced. (Taking the liquid in cell
// order instead, so that the lanes of one cell stream the same candidates, was measured slower: 1.27 against
// 1.15 ms per build on the 512-world bowl pack, the scattered 2-byte row writes costing more than the shared loads
// save.)
template <bool WIDE, int BB, bool SUP>
__global__ void __launch_bounds__(BB) k_build_lists(Ctx c) {
    using RowT = typename std::conditional<WIDE, int, uint16_t>::type;
    __shared__ RowT row[ROW_BUF * BB];
    int i = xcd_particle();
    if (i >= c.n) return;
    int g = c.group[i];
    if (SUP) {
        int w = c.world[i];
        if (w < 0 || c.skin_flag[w] == 0) return;
        c.q_ref[i] = c.q4[i];
        c.g_ref[i] = g;
        if (g < 0) {
            c.sup_count[i] = 0;
            return;
        }
    } else if (g < 0) {
        c.count[i] = 0;
        return;
    }
    RowT* L = row + threadIdx.x;
    f3 xi = xyz(c.q4[i]);
    int nn, total;
    if (!SUP) {
        build_row<WIDE, BB>(c, i, g, xi, c.h * c.h, c.cap, L, nn, total);
        write_row<WIDE, BB>(c, i, g, nn, total, L);
        return;
    }
    float hs = c.h + c.skin;
    build_row<WIDE, BB>(c, i, g, xi, hs * hs, c.cap_s, L, nn, total);
    if (total <= c.cap_s) {
        for (int k = 0; k < nn; ++k) {
            int v = L[k * BB];
            c.sup[int64_t(k) * c.n + i] = uint16_t(v & 65535);
            if (WIDE) c.sup_hi[int64_t(k) * c.n + i] = uint8_t(v >> 16);
        }
        c.sup_count[i] = nn;
        return;
    }
    build_row<WIDE, BB>(c, i, g, xi, c.h * c.h, c.cap, L, nn, total);
    write_row<WIDE, BB>(c, i, g, nn, total, L);
    c.sup_count[i] = -1;
    c.skin_force[g] = 1;
}

// LDS-staged list build (DFSPH_OPT_BUILD_STAGED, plain rows). A block takes BB consecutive particles of one chunk
// (one world), finds the box of grid cells within one cell of its liquid particles' cells, and copies the cell-ordered
// records of that box (q4s, contiguous per z-column of cells) into LDS once, with coalesced loads; each lane then
// walks its 27 cells over the staged records. The lanes of a block share most of their cells, so each record crosses
// the vector memory path once per block instead of once per lane that visits it. A box with more records than the
// staging array is staged in z-slabs, each lane visiting the part of its cells inside each slab: a row is the lowest
// `cap` of all accepted ids, sorted (build_row's overflow step keeps the lowest `cap` so far), which does not depend
// on the visit order, so rows, counts and overflow are those of k_build_lists bit for bit. A box whose cell table or
// single z-layer does not fit falls back to the per-lane walk from global memory.
// Measured on the 512-world bowl pack: 6.1 ms per build against the per-lane walk's 1.15 ms. The staged records and
// the per-lane row buffers (which collect each row before its sort) hold 37 KB of LDS per 64-thread block, so a CU
// keeps 3.6 waves resident against the per-lane walk's 16, and a lone wave per SIMD cannot cover its own dependent
// LDS and ALU latencies (the VALU is 17 % busy, waves wait 52 % of their cycles). The option stays off by default.
constexpr int ST_REC = 16;      // staged records per thread of the block
constexpr int ST_CELLS = 2048;  // box cell-table entries (each column's z cells plus its end)
constexpr int ST_COLS = 256;    // box z-columns
constexpr int ST_LAYERS = 128;  // z-layers of a box
constexpr int ST_ILP = 8;       // staging loads in flight per thread
// exclusive prefix of v over the block (BB a multiple of 64, at most 1024); *tot the block total
template <int BB>
__device__ __forceinline__ int block_prefix(int v, int* tot, int* wsum) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int d = 1; d < 64; d <<= 1) {
        int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    int off = 0, all = 0;
    for (int k = 0; k < BB / 64; ++k) {
        if (k < wid) off += wsum[k];
        all += wsum[k];
    }
    *tot = all;
    __syncthreads();
    return off + x - v;
}
template <bool WIDE, int BB>
__global__ void __launch_bounds__(BB) k_build_lists_staged(Ctx c) {
    using RowT = typename std::conditional<WIDE, int, uint16_t>::type;
    constexpr int SUB = CHUNK / BB, NREC = ST_REC * BB, NCELL = ST_CELLS, NCOL = ST_COLS, CPT = NCOL / BB;
    __shared__ RowT row[ROW_BUF * BB];
    __shared__ float4 srec[NREC + BUILD_U];  // + one candidate batch read past a range's end
    __shared__ int scell[NCELL];             // cell_start of the box's cells: column-major, z0 .. z1 + 1 per column
    __shared__ int coloff[NCOL + 1];         // staged offset of each column's slab segment
    __shared__ int ltot[ST_LAYERS];          // records per z-layer of the box
    __shared__ int sbeg[ST_LAYERS + 1];      // z-slab starts (box layers)
    __shared__ int bx[8];                    // min x, y, z, max x, y, z, slab count, fallback
    __shared__ int wsum[BB / 64];
    static_assert(CHUNK % BB == 0 && BB % 64 == 0 && NCOL % BB == 0, "whole waves dividing a chunk");
    const int t = int(threadIdx.x);
    const int bq = xcd_block(int(blockIdx.x), int(gridDim.x));
    const int ch = bq / SUB, i0 = c.ch_begin[ch] + (bq % SUB) * BB;
    if (i0 >= c.ch_end[ch]) return;  // the sub-block past a world's last particle (uniform)
    const int i = i0 + t, w = c.ch_world[ch];
    const bool valid = i < c.ch_end[ch];
    const int g = valid ? c.group[i] : -1;
    if (valid && g < 0) c.count[i] = 0;
    const bool mine = valid && g >= 0;  // liquid: g is the world
    f3 xi{0.0f, 0.0f, 0.0f};
    int cx = 0, cy = 0, cz = 0;
    if (mine) {
        xi = xyz(c.q4[i]);
        cell_of(c, w, xi, cx, cy, cz);
    }
    if (t < 8) bx[t] = t < 3 ? INT_MAX : (t < 6 ? -1 : 0);
    __syncthreads();
    if (mine) {
        atomicMin(&bx[0], cx);
        atomicMin(&bx[1], cy);
        atomicMin(&bx[2], cz);
        atomicMax(&bx[3], cx);
        atomicMax(&bx[4], cy);
        atomicMax(&bx[5], cz);
    }
    __syncthreads();
    if (bx[3] < 0) return;  // no liquid in the block (uniform)
    const int* D = c.g_dims + 3 * w;
    const int x0 = max(bx[0] - 1, 0), y0 = max(bx[1] - 1, 0), z0 = max(bx[2] - 1, 0);
    const int nbx = min(bx[3] + 1, D[0] - 1) - x0 + 1, nby = min(bx[4] + 1, D[1] - 1) - y0 + 1;
    const int nbz = min(bx[5] + 1, D[2] - 1) - z0 + 1, ncz = nbz + 1, ncol = nbx * nby;
    const float h2 = c.h * c.h;
    const int cap = c.cap, base = c.ws[w];
    RowT* L = row + t;
    int nn = 0, total = 0;
    const bool fits = ncol <= NCOL && ncol * ncz <= NCELL && nbz <= ST_LAYERS;  // uniform
    if (fits) {
        const int cb = c.cell_base[w];
        for (int k0 = t; k0 < ncol * ncz; k0 += ST_ILP * BB) {  // ST_ILP independent loads in flight per thread
            int v[ST_ILP];
#pragma unroll
            for (int u = 0; u < ST_ILP; ++u) {
                const int k = min(k0 + u * BB, ncol * ncz - 1);
                const int col = k / ncz, dz = k - col * ncz;
                v[u] = c.cell_start[cb + ((x0 + col / nby) * D[1] + y0 + col % nby) * D[2] + z0 + dz];
            }
#pragma unroll
            for (int u = 0; u < ST_ILP; ++u)
                if (k0 + u * BB < ncol * ncz) scell[k0 + u * BB] = v[u];
        }
        for (int k = t; k < nbz; k += BB) ltot[k] = 0;
        __syncthreads();
        for (int k = t; k < ncol * nbz; k += BB) {
            const int col = k / nbz, dz = k - col * nbz;
            atomicAdd(&ltot[dz], scell[col * ncz + dz + 1] - scell[col * ncz + dz]);
        }
        __syncthreads();
        if (t == 0) {  // greedy z-slabs of at most NREC records; a layer over NREC on its own: fall back
            int ns = 0, acc = 0;
            sbeg[0] = 0;
            for (int dz = 0; dz < nbz; ++dz) {
                if (ltot[dz] > NREC) bx[7] = 1;
                if (acc > 0 && acc + ltot[dz] > NREC) {
                    sbeg[++ns] = dz;
                    acc = 0;
                }
                acc += ltot[dz];
            }
            sbeg[++ns] = nbz;
            bx[6] = ns;
        }
        __syncthreads();
    }
    if (!fits || bx[7] != 0) {  // the per-lane walk over global memory (uniform)
        if (mine) {
            build_row<WIDE, BB>(c, i, g, xi, h2, cap, L, nn, total);
            write_row<WIDE, BB>(c, i, g, nn, total, L);
        }
        return;
    }
    const int nslab = bx[6];
    for (int sl = 0; sl < nslab; ++sl) {
        const int zs = sbeg[sl], ze = sbeg[sl + 1];  // box layers [zs, ze)
        // the slab segment of every column, prefix-summed: CPT columns per thread
        int v[CPT], sum = 0;
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int col = t * CPT + u;
            v[u] = col < ncol ? scell[col * ncz + ze] - scell[col * ncz + zs] : 0;
            sum += v[u];
        }
        int all;
        int off = block_prefix<BB>(sum, &all, wsum);
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int col = t * CPT + u;
            if (col < ncol) coloff[col] = off;
            off += v[u];
        }
        if (t == 0) coloff[ncol] = all;
        __syncthreads();
        for (int r0 = t; r0 < all; r0 += ST_ILP * BB) {  // ST_ILP records per thread per round, loads in flight together
            int src[ST_ILP];
#pragma unroll
            for (int u = 0; u < ST_ILP; ++u) {  // the column holding staged record r: the last with coloff <= r
                const int r = min(r0 + u * BB, all - 1);
                int lo = 0, hi = ncol - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (coloff[mid] <= r) lo = mid;
                    else hi = mid - 1;
                }
                src[u] = scell[lo * ncz + zs] + (r - coloff[lo]);
            }
            float4 rec[ST_ILP];
#pragma unroll
            for (int u = 0; u < ST_ILP; ++u) rec[u] = c.q4s[src[u]];
#pragma unroll
            for (int u = 0; u < ST_ILP; ++u)
                if (r0 + u * BB < all) srec[r0 + u * BB] = rec[u];
        }
        __syncthreads();
        if (mine) {
            for (int dx = -1; dx <= 1; ++dx) {
                const int X = cx + dx;
                if (X < 0 || X >= D[0]) continue;
                for (int dy = -1; dy <= 1; ++dy) {
                    const int Y = cy + dy;
                    if (Y < 0 || Y >= D[1]) continue;
                    const int col = (X - x0) * nby + (Y - y0);
                    const int za = max(max(cz - 1, 0) - z0, zs), zb = min(min(cz + 1, D[2] - 1) - z0, ze - 1);
                    if (za > zb) continue;
                    const int* sc = scell + col * ncz;
                    const int e0 = coloff[col] + sc[za] - sc[zs], e1 = coloff[col] + sc[zb + 1] - sc[zs];
                    for (int e = e0; e < e1; e += BUILD_U) {
                        if (nn > ROW_BUF - BUILD_U) {  // room for a whole batch: keep the lowest cap
                            sort_row<RowT, BB>(L, nn);
                            nn = cap;
                        }
                        float4 pc[BUILD_U];
#pragma unroll
                        for (int u = 0; u < BUILD_U; ++u) pc[u] = srec[e + u];
#pragma unroll
                        for (int u = 0; u < BUILD_U; ++u) {
                            const int j = __float_as_int(pc[u].w);
                            const f3 dd = xyz(pc[u]) - xi;
                            const bool ok = (e + u < e1) & (j != i) & (dot(dd, dd) < h2);
                            if (ok) {
                                total += 1;
                                L[nn * BB] = RowT(j - base);
                                nn += 1;
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();  // the next slab overwrites the staged records
    }
    if (mine) {
        sort_row<RowT, BB>(L, nn);
        nn = nn < cap ? nn : cap;
        write_row<WIDE, BB>(c, i, g, nn, total, L);
    }
}

// neighbors_skin.py:34-42 seed: this step's flags start from the forced worlds
__global__ void k_skin_seed(Ctx c) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W) return;
    c.skin_flag[w] = c.skin_force[w];
    c.skin_force[w] = 0;
}

// neighbors_skin.py:45-60 check: a world rebuilds when one of its particles changed group or moved more than
// 0.49 skin since its last build (a NaN move counts)
__global__ void __launch_bounds__(256) k_skin_check(Ctx c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    int w = c.world[i];
    if (w < 0) return;
    int g = c.group[i];
    bool moved = g != c.g_ref[i];
    if (!moved && g >= 0) {
        f3 d = xyz(c.q4[i]) - xyz(c.q_ref[i]);
        float half = 0.49f * c.skin;
        moved = !(dot(d, d) <= half * half);
    }
    if (moved) c.skin_flag[w] = 1;
}

// neighbors_skin.py:115-145 filter: the superset entries now within h, in the row's ascending order, the first
// `cap` of them, overflow counted: the rows, counts and overflow of a plain build at these positions
template <bool WIDE>
__global__ void __launch_bounds__(256) k_skin_filter(Ctx c) {
    int i = xcd_particle();
    if (i >= c.n) return;
    int g = c.group[i];
    if (g < 0) {
        c.count[i] = 0;
        return;
    }
    int ns = c.sup_count[i];
    if (ns < 0) return;  // a plain row written by this step's build
    int base = c.ws[g];
    f3 xi = xyz(c.q4[i]);
    float h2 = c.h * c.h;
    int n = 0, total = 0;
    for (int k = 0; k < ns; ++k) {
        int jl = int(c.sup[int64_t(k) * c.n + i]);
        if (WIDE) jl += int(c.sup_hi[int64_t(k) * c.n + i]) << 16;
        f3 dd = xyz(c.q4[base + jl]) - xi;
        if (dot(dd, dd) < h2) {
            total += 1;
            if (n < c.cap) {
                c.nbr[row_at(c, i, n)] = uint16_t(jl & 65535);
                if (WIDE) c.nbr_hi[row_at(c, i, n)] = uint8_t(jl >> 16);
                n += 1;
            }
        }
    }
    if (total > c.cap) atomicAdd(c.st_overflow + g, 1);
    c.count[i] = n;
}

// ----------------------------------------------------------------------------------------------- particle walls
// particle_walls.py:65-84: colour, grad colour and mean velocity of world w's wall particles at x
constexpr int PW_ROW = 32;  // wall particles recorded per liquid particle for the reaction split

// rec >= 0: record the visited wall particles of liquid particle rec (wr_id / wr_w / wr_n)
constexpr int PW_U = 4;  // wall candidates loaded together in pw_sample (q4s holds BUILD_U >= PW_U padding records)
static_assert(BUILD_U >= PW_U, "pw_sample reads its batches unclamped");
__device__ void pw_sample(const Ctx& c, f3 x, int w, float& col, f3& gc, f3& mv, int rec) {
    const KC K = make_kc(c.h, c.kind);
    col = 0.0f;
    gc = f3{0.0f, 0.0f, 0.0f};
    mv = f3{0.0f, 0.0f, 0.0f};
    if (c.NW == 0) return;
    int g = c.W + w, cx, cy, cz;
    int own = cell_of(c, g, x, cx, cy, cz);
    if (!c.wall_near[own - c.wall_first]) {  // no wall particle in the 27 cells: the sums below stay 0
        if (rec >= 0) c.wr_n[rec] = 0;
        return;
    }
    int nrec = 0;
    const int* d = c.g_dims + 3 * w;
    // |d|^2 > 1.0001 h^2 (rounded either way) puts |d| above h (1 + 4e-5) and its rounded quotient by h above 1, where
    // the kernel is 0: such a candidate is skipped before the kernel's square root and quotient
    const float out2 = 1.0001f * (c.h * c.h);
    // the 27 cells in the order z, y, x (x fastest), r = 9 (dz + 1) + 3 (dy + 1) + (dx + 1); the range of cell r + 1
    // is loaded before cell r's candidates (an empty range off the grid)
    auto cell_range = [&](int r, int& a0, int& a1) {
        const int X = cx + r % 3 - 1, Y = cy + (r / 3) % 3 - 1, Z = cz + r / 9 - 1;
        a0 = a1 = 0;
        if (X < 0 || Y < 0 || Z < 0 || X >= d[0] || Y >= d[1] || Z >= d[2]) return;
        const int cc = c.cell_base[g] + (X * d[1] + Y) * d[2] + Z;
        a0 = c.cell_start[cc];
        a1 = c.cell_start[cc + 1];
    };
    int n0, n1;
    auto walk = [&](auto kv) {  // the kernel variant chosen once for the walk (kernel_w_rv)
        constexpr int KV = decltype(kv)::value;
        cell_range(0, n0, n1);
        for (int r = 0; r < 27; ++r) {
            const int b0 = n0, e1 = n1;
            if (r < 26) cell_range(r + 1, n0, n1);
            // q4s: the grid's cell-ordered copy of q4 with the particle id in w (the same positions; one
            // contiguous load per candidate, PW_U of them in flight, read unclamped from the padded array, then
            // taken in order). Outside the support W = 0 makes V_j W zero, so V_j is read for the candidates
            // within h only.
            for (int e0 = b0; e0 < e1; e0 += PW_U) {
                float4 pb[PW_U];
    #pragma unroll
                for (int u = 0; u < PW_U; ++u) pb[u] = c.q4s[e0 + u];
    #pragma unroll
                for (int u = 0; u < PW_U; ++u) {
                    if (e0 + u >= e1) continue;
                    const float4 p4 = pb[u];
                    f3 dd = x - f3{p4.x, p4.y, p4.z};
                    if (dot(dd, dd) > out2) continue;  // kernel_w is 0 there
                    float kw = kernel_w_rv<KV>(dd, K);
                    if (!(kw > 0.0f)) continue;
                    int j = __float_as_int(p4.w);
                    float vj = c.wvol[j];
                    float wj = vj * kw;
                    if (wj > 0.0f) {
                        col += wj;
                        gc = gc + vj * pair_grad(dd, kernel_grad_scale_v<KV>(dd, K));
                        mv = mv + wj * V(c, j);
                        if (rec >= 0 && nrec < PW_ROW) {
                            c.wr_id[int64_t(nrec) * c.n + rec] = c.wcidx[j];
                            c.wr_w[int64_t(nrec) * c.n + rec] = wj;
                        }
                        nrec += 1;
                    }
                }
            }
        }
    };
    if (K.fast && K.kind == 0) walk(std::integral_constant<int, 0>{});
    else walk(std::integral_constant<int, 1>{});
    if (rec >= 0) c.wr_n[rec] = nrec <= PW_ROW ? nrec : -1;
    if (col > 0.0f) mv = mv / col;
}
// particle_walls.py:87-98
__device__ float colour_distance(const Ctx& c, float col) {
    float lo = -c.h, hi = c.h;
    for (int k = 0; k < 30; ++k) {
        float mid = 0.5f * (lo + hi);
        float psi, dpsi;
        halfspace_volume(mid, c.h, 0.0f, c.kind, psi, dpsi);
        if (psi > col) lo = mid;
        else hi = mid;
    }
    return 0.5f * (lo + hi);
}

// particle_walls.py:138-177 _accumulate_kernel: each particle-wall slot's impulse J, split over the world's wall
// particles by V_j W_ij / c_i in int64 at 2^50 (the rounding remainder to the largest share, so the parts sum exactly
// to the slot's fixed-point impulse), and the walls' work -J.v_b
template <class AddF>
__device__ __forceinline__ void pw_accumulate_one(const Ctx& c, int i, AddF add) {
    const KC K = make_kc(c.h, c.kind);
    const double IMP = 1125899906842624.0;
    int64_t b4 = 4 * int64_t(i);
    for (int k = 0; k < 4; ++k) {
        if (c.s_shape[b4 + k] != c.S) continue;
        f3 J = ld3(c.s_imp, b4 + k);
        if (J.x == 0.0f && J.y == 0.0f && J.z == 0.0f) continue;
        int w = c.world[i];
        atomicAdd(reinterpret_cast<unsigned long long*>(c.pw_work) + w,
                  static_cast<unsigned long long>(llround(double(-dot(J, ld3(c.s_vb, b4 + k))) * IMP)));
        if (c.one_way) continue;
        f3 x = xyz(c.q4[i]);
        float col = c.pw_col[i];  // pw_sample at the same positions in this step's k_slots
        if (col <= 0.0f) continue;
        long long left[3] = {llround(double(J.x) * IMP), llround(double(J.y) * IMP), llround(double(J.z) * IMP)};
        int best = -1;
        float sbest = 0.0f;
        auto share = [&](int cj, float wj) {  // wj = V_j W_ij as the slot pass summed it
            float sj = wj / col;
            if (sj > 0.0f) {
                long long part[3] = {llround(double(J.x * sj) * IMP), llround(double(J.y * sj) * IMP),
                                     llround(double(J.z * sj) * IMP)};
                add(cj, part);
                for (int q = 0; q < 3; ++q) left[q] -= part[q];
                if (sj > sbest) {
                    sbest = sj;
                    best = cj;
                }
            }
        };
        const int nrow = c.wr_n[i];
        if (nrow >= 0) {  // the slot pass's wall particles, in its visiting order
            for (int k = 0; k < nrow; ++k) share(c.wr_id[int64_t(k) * c.n + i], c.wr_w[int64_t(k) * c.n + i]);
        } else {  // more than the row holds: visit the cells again, in the same order
            int g = c.W + w, cx, cy, cz;
            cell_of(c, g, x, cx, cy, cz);
            const int* d = c.g_dims + 3 * w;
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int X = cx + dx, Y = cy + dy, Z = cz + dz;
                        if (X < 0 || Y < 0 || Z < 0 || X >= d[0] || Y >= d[1] || Z >= d[2]) continue;
                        int cc = c.cell_base[g] + (X * d[1] + Y) * d[2] + Z;
                        for (int e = c.cell_start[cc]; e < c.cell_start[cc + 1]; ++e) {
                            int j = c.items[e];
                            float wj = c.wvol[j] * kernel_w(dist(x - xyz(c.q4[j]), K), K);
                            if (wj > 0.0f) share(c.wcidx[j], wj);
                        }
                    }
        }
        if (best >= 0) add(best, left);
    }
}
// The wall particles' shares go through a per-block table in LDS (open addressing on the compact wall index, int64
// sums), flushed with one global atomic per touched wall particle and component: a block's liquid pushes on a few
// hundred wall particles many times over. Integer sums do not depend on their order, so the totals are those of
// one global atomic per share; a share whose probe finds no free slot goes to global memory directly.
constexpr int EXCH_T = 1024;  // slots of the block's table (28 KB of LDS)
__global__ void __launch_bounds__(256) k_pw_accumulate(Ctx c) {
    __shared__ int tkey[EXCH_T];
    __shared__ unsigned long long tval[3 * EXCH_T];
    for (int t = threadIdx.x; t < EXCH_T; t += blockDim.x) {
        tkey[t] = -1;
        tval[3 * t] = tval[3 * t + 1] = tval[3 * t + 2] = 0ull;
    }
    __syncthreads();
    unsigned long long* gx = reinterpret_cast<unsigned long long*>(c.pw_exch);
    const int i = xcd_particle();
    if (i < c.n)
        pw_accumulate_one(c, i, [&](int cj, const long long (&v)[3]) {
            const unsigned h = (unsigned(cj) * 2654435761u) >> 22;  // 10 bits
            for (int p = 0; p < 16; ++p) {
                const int t = int((h + unsigned(p)) & unsigned(EXCH_T - 1));
                const int k = atomicCAS(&tkey[t], -1, cj);
                if (k == -1 || k == cj) {
                    for (int q = 0; q < 3; ++q) atomicAdd(&tval[3 * t + q], static_cast<unsigned long long>(v[q]));
                    return;
                }
            }
            for (int q = 0; q < 3; ++q) atomicAdd(gx + 3 * int64_t(cj) + q, static_cast<unsigned long long>(v[q]));
        });
    __syncthreads();
    for (int t = threadIdx.x; t < EXCH_T; t += blockDim.x) {
        const int cj = tkey[t];
        if (cj < 0) continue;
        for (int q = 0; q < 3; ++q)
            if (tval[3 * t + q] != 0ull) atomicAdd(gx + 3 * int64_t(cj) + q, tval[3 * t + q]);
    }
}



// ----------------------------------------------------------------------------------------------- boundary slots
struct SlotRow {
    int shape[4];
    float phi[4];
    f3 nrm[4];
};

// Element k of a 4-array with a run-time k, by selects over the constant indices, so the array stays in registers.
template <typename T>
__device__ __forceinline__ T pick4(const T (&a)[4], int k) {
    T r = a[0];
#pragma unroll
    for (int m = 1; m < 4; ++m)
        if (k == m) r = a[m];
    return r;
}
template <typename T>
__device__ __forceinline__ void put4(T (&a)[4], int k, T v) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
        if (k == m) a[m] = v;
}

// boundary_slots.py:13-38 (slot normals are not shifted, as in the reference). The scan runs from slot nn - 1 down
// and stops at the first phi <= the new one; the shifts run from the top, every index a constant.
__device__ __forceinline__ int slot_insert(SlotRow& R, int nn, int s, float phi) {
    int pos = nn;
    bool scan = true;
#pragma unroll
    for (int k = 3; k >= 0; --k)
        if (k < nn && scan) {
            if (R.phi[k] <= phi) scan = false;
            else pos = k;
        }
    if (pos >= 4) return -1;
    int last = nn >= 4 ? 3 : nn;
#pragma unroll
    for (int k = 3; k >= 1; --k)
        if (k <= last && k > pos) {
            R.shape[k] = R.shape[k - 1];
            R.phi[k] = R.phi[k - 1];
        }
    put4(R.shape, pos, s);
    put4(R.phi, pos, phi);
    return pos;
}
__device__ __forceinline__ void slot_copy(SlotRow& R, int src, int dst) {
    R.shape[dst] = R.shape[src];
    R.phi[dst] = R.phi[src];
    R.nrm[dst] = R.nrm[src];
}
// boundary_slots.py:50-102
__device__ int shape_insert(SlotRow& R, int nn, int s, float phi, f3 nrm, float near_) {
    if (nn < 4) {
        int rr = slot_insert(R, nn, s, phi);
        if (rr >= 0) {
#pragma unroll
            for (int k = 3; k >= 1; --k)
                if (k <= nn && k > rr) R.nrm[k] = R.nrm[k - 1];
            put4(R.nrm, rr, nrm);
            return 1;
        }
        return 0;
    }
    const float cos_same = 0.85f;
    int ev = -1;
    float ev_phi = -1.0f;
    for (int m = 0; m < 4; ++m)
        if (phi >= near_ && R.phi[m] <= phi && dot(R.nrm[m], nrm) > cos_same) {
            ev = 4;
            ev_phi = phi;
        }
    for (int k = 0; k < 4; ++k) {
        float pk = R.phi[k];
        int dup = 0;
        for (int m = 0; m < 4; ++m)
            if (m != k && R.phi[m] <= pk && dot(R.nrm[m], R.nrm[k]) > cos_same) dup = 1;
        if (phi <= pk && dot(nrm, R.nrm[k]) > cos_same) dup = 1;
        if (dup == 1 && pk >= near_ && pk > ev_phi) {
            ev = k;
            ev_phi = pk;
        }
    }
    if (ev < 0 && phi < R.phi[3]) ev = 3;
    if (ev < 0 || ev == 4) return 0;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (k >= ev) slot_copy(R, k + 1, k);
    int pos = 3;
    bool scan = true;
#pragma unroll
    for (int k = 3; k >= 1; --k)
        if (scan && R.phi[k - 1] > phi) {
            slot_copy(R, k - 1, k);
            pos = k - 1;
        } else {
            scan = false;
        }
    put4(R.shape, pos, s);
    put4(R.phi, pos, phi);
    put4(R.nrm, pos, nrm);
    return 1;
}

__device__ __forceinline__ void taper(const Ctx& c, float& v0, float& v1, float phi) {
    if (phi <= c.taper0) return;
    float L = c.taper1 - c.taper0;
    float t = fmn((phi - c.taper0) / L, 1.0f);
    float sm = 1.0f - t * t * (3.0f - 2.0f * t);
    float dsm = 6.0f * t * (t - 1.0f) / L;
    float a = v0 * sm, b = v1 * sm + v0 * dsm;
    v0 = a;
    v1 = b;
}

// boundary.py:190-238 (find) and 241-396 (fill, inclusion-exclusion) for one fluid particle
__global__ void __launch_bounds__(256) k_slots(Ctx c) {
    int i = xcd_particle();
    if (i >= c.n) return;
    SlotRow R;
    for (int k = 0; k < 4; ++k) {
        R.shape[k] = -1;
        R.phi[k] = 0.0f;
        R.nrm[k] = f3{0.0f, 0.0f, 0.0f};
    }
    int64_t b4 = 4 * int64_t(i);
    if (!is_fluid(c, i)) {
        // A row whose four slots are empty is already clear: its psi, grad psi and impulses are zero (the fill
        // writes zeros for empty slots, impulses go only to slots with a wall) and nothing writes the row of a
        // particle that is not liquid. Only a particle that left the liquid with a slot in use is cleared.
        const int4 sh = *reinterpret_cast<const int4*>(c.s_shape + b4);
        if (sh.x == -1 && sh.y == -1 && sh.z == -1 && sh.w == -1) return;
        for (int k = 0; k < 4; ++k) {
            st3(c.s_imp, b4 + k, f3{0.0f, 0.0f, 0.0f});
            c.s_shape[b4 + k] = -1;
            c.s_psi[b4 + k] = 0.0f;
            st3(c.s_gpsi, b4 + k, f3{0.0f, 0.0f, 0.0f});
        }
        return;
    }
    // A row whose four slots were empty is clear (as above): its impulses are still zero, so _begin_step's clear
    // (solver.py:289-291) is needed only for a row that had a slot, and a row that stays empty is not written again.
    const int4 sh_old = *reinterpret_cast<const int4*>(c.s_shape + b4);
    const bool was_empty = sh_old.x == -1 && sh_old.y == -1 && sh_old.z == -1 && sh_old.w == -1;
    if (!was_empty)
        for (int k = 0; k < 4; ++k) st3(c.s_imp, b4 + k, f3{0.0f, 0.0f, 0.0f});
    int w = c.world[i];
    f3 x = xyz(c.q4[i]);
    int nn = 0;
    for (int e = c.sh_start[w]; e < c.sh_start[w + 1]; ++e) {
        int s = c.sh_ids[e];
        Xf X = load_xf(c.sh_x + 7 * s);
        int geo = c.sh_geo[s];
        if (geo != GEO_PLANE)
            if (length(x - X.p) > c.sh_radius[s] + c.h) continue;
        f3 xl = xf_inv_point(X, x);
        float phi;
        f3 gl;
        shape_sdf(geo, ld3(c.sh_scale, s), xl, phi, gl);
        if (phi < c.reach && phi + c.offset < c.h) {
            f3 nw = normalize(xf_vector(X, gl));
            if (shape_insert(R, nn, s, phi, nw, 0.1f * c.h) > 0 && nn < 4) nn += 1;
        }
    }
    float pcol = 0.0f;
    f3 pgc{0.0f, 0.0f, 0.0f}, pmv{0.0f, 0.0f, 0.0f};
    bool have_pw = false;
    float pw_phi = 1.0e30f;
    if (c.NW > 0) {
        pw_sample(c, x, w, pcol, pgc, pmv, i);
        c.pw_col[i] = pcol;
        have_pw = true;
        if (!(pcol < 1.0e-4f || length(pgc) <= 0.0f)) pw_phi = colour_distance(c, fmn(pcol, 1.0f));
        const float phi = pw_phi;
        if (phi < c.reach && phi + c.offset < c.h)
            if (slot_insert(R, nn, c.S, phi) >= 0 && nn < 4) nn += 1;
    }
    if (nn == 0 && was_empty) return;  // the row stays empty: its shapes, psi and grad psi are stored already
    // fill (boundary.py:290-333)
    float psi[4];
    f3 gp[4];
    for (int k = 0; k < 4; ++k) {
        int s = pick4(R.shape, k);
        c.s_shape[b4 + k] = s;
        if (s < 0) {
            put4(psi, k, 0.0f);
            put4(gp, k, f3{0.0f, 0.0f, 0.0f});
            continue;
        }
        if (s == c.S && have_pw) {  // particle wall: its local plane (boundary.py:272-287)
            float lg = length(pgc);
            f3 nrm{0.0f, 0.0f, 1.0f};
            if (lg > 0.0f) nrm = (-pgc) / lg;
            // the insert's distance: the slot exists only if it was computed, and pcol >= 0 (a sum of V_j W > 0)
            // makes fcl(pcol, 0, 1) = fmn(pcol, 1)
            float phi = pw_phi;
            f3 xb = x - phi * nrm;
            float v0, v1;
            halfspace_volume(phi + c.offset, c.h, 0.0f, c.kind, v0, v1);
            v0 = v0 * c.amplitude;
            v1 = v1 * c.amplitude;
            taper(c, v0, v1, phi);
            c.s_phi[b4 + k] = phi;
            put4(psi, k, v0);
            put4(gp, k, v1 * nrm);
            st3(c.s_xb, b4 + k, xb);
            st3(c.s_vb, b4 + k, pmv);
            st3(c.s_xbl, b4 + k, xb);
            st3(c.s_nbl, b4 + k, nrm);
            continue;
        }
        Xf X = load_xf(c.sh_x + 7 * s);
        int geo = c.sh_geo[s];
        f3 sc = ld3(c.sh_scale, s);
        f3 xl = xf_inv_point(X, x);
        float phi;
        f3 gl;
        shape_sdf(geo, sc, xl, phi, gl);
        f3 nrm = normalize(xf_vector(X, gl));
        f3 nl = normalize(gl);
        st3(c.s_nbl, b4 + k, nl);
        st3(c.s_xbl, b4 + k, xl - phi * nl);
        float H = 0.0f;
        if (c.curvature > 0 && geo == GEO_SPHERE) {
            float kmax = 2.0f / c.h;
            H = -2.0f * fcl(1.0f / sc.x, -kmax, kmax);
        }
        float v0, v1;
        halfspace_volume(phi + c.offset, c.h, H, c.kind, v0, v1);
        v0 = v0 * c.amplitude;
        v1 = v1 * c.amplitude;
        taper(c, v0, v1, phi);
        c.s_phi[b4 + k] = phi;
        put4(psi, k, v0);
        put4(gp, k, v1 * nrm);
        st3(c.s_xb, b4 + k, x - phi * nrm);
        st3(c.s_vb, b4 + k, f3{0.0f, 0.0f, 0.0f});
    }
    // inclusion-exclusion (boundary.py:370-396); the gradients of earlier slots are read after their update
    float ps[4] = {psi[0], psi[1], psi[2], psi[3]};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        if (R.shape[b] < 0) continue;
        f3 gb = gp[b];
        float ob = 0.0f, sb = 0.0f;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
            if (cc == b || R.shape[cc] < 0) continue;
            f3 gcv = gp[cc];
            float lb = length(gb), lc = length(gcv);
            float cbc = 0.0f;
            if (lb > 0.0f && lc > 0.0f) cbc = fcl(dot(gb, gcv) / (lb * lc), -1.0f, 1.0f);
            ob += (1.0f - fabsf(cbc)) * ps[cc];
            if (cbc > 0.0f && (ps[cc] > ps[b] || (ps[cc] == ps[b] && cc < b))) sb += cbc;
        }
        ob = fmn(ob, 1.0f);
        psi[b] = ps[b] * fmx(1.0f - 0.5f * ob - sb, 0.0f);
        gp[b] = gb * fmx(1.0f - ob - sb, 0.0f);
    }
    for (int k = 0; k < 4; ++k) {
        c.s_psi[b4 + k] = psi[k];
        st3(c.s_gpsi, b4 + k, gp[k]);
    }
}

// ----------------------------------------------------------------------------------------------- density / alpha
#define CHUNK_PROLOGUE                                     \
    int b_ = xcd_block(int(blockIdx.x), int(gridDim.x));   \
    int i = c.ch_begin[b_] + threadIdx.x; \
    bool valid = i < c.ch_end[b_];

// ----------------------------------------------------------------------------------------------- staged halos
// The gather passes of the pressure solves read, per pair, a 16-32 B record of the neighbour. Through the vector
// memory path that is one gather instruction per record and about 20 distinct cache lines per instruction. A chunk's
// 256 particles share most of their neighbours (their halo is about 4 chunks of records), so the staged passes load
// the halo's records into LDS once per launch, with loads that walk ascending ids, and gather from LDS by each row
// entry's halo index. The neighbour order of every sum is the row's, so results are bit for bit those of the direct
// passes.
constexpr int HALO_MAX = 1536;   // records a chunk's halo may hold in LDS (48 KB of {q4, acc4} in k_apply_h)
constexpr int HALO_BITS = 32768; // world-local id range a halo may span (the bitmap of k_halo_build, 4 KB of LDS)

// One block per chunk: the distinct world-local ids of the chunk's row entries by a bitmap over their id range, in
// ascending order (h_ids, as global ids), and each entry's rank among them (h_slot). A chunk whose entries span more
// than HALO_BITS ids or hold more than HALO_MAX distinct neighbours gets h_count -1: the passes gather its
// neighbours directly.
template <bool WIDE>
__global__ void __launch_bounds__(256) k_halo_build(Ctx c) {
    __shared__ unsigned bits[HALO_BITS / 32];
    __shared__ int wpre[HALO_BITS / 32];
    __shared__ int red[8];
    CHUNK_PROLOGUE
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int cnt = valid ? c.count[i] : 0;
    const int w = c.ch_world[b_];
    auto entry = [&](int k) {
        const int64_t r = row_at(c, i, k);
        int e = int(c.nbr[r]);
        if (WIDE) e += int(c.nbr_hi[r]) << 16;
        return e;
    };
    int lo = INT_MAX, hi = -1;
    for (int k = 0; k < cnt; ++k) {
        int e = entry(k);
        lo = min(lo, e);
        hi = max(hi, e);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d, 64));
        hi = max(hi, __shfl_xor(hi, d, 64));
    }
    if (lane == 0) {
        red[wv] = lo;
        red[4 + wv] = hi;
    }
    __syncthreads();
    lo = min(min(red[0], red[1]), min(red[2], red[3]));
    hi = max(max(red[4], red[5]), max(red[6], red[7]));
    if (hi < 0) {  // no row entries: an empty halo
        if (t == 0) c.h_count[b_] = 0;
        return;
    }
    if (hi - lo + 1 > HALO_BITS) {
        if (t == 0) c.h_count[b_] = -1;
        return;
    }
    const int nw = (hi - lo + 1 + 31) >> 5;
    for (int q = t; q < nw; q += 256) bits[q] = 0u;
    __syncthreads();
    for (int k = 0; k < cnt; ++k) {
        int d = entry(k) - lo;
        atomicOr(&bits[d >> 5], 1u << (d & 31));
    }
    __syncthreads();
    // exclusive prefix of the words' popcounts: thread t owns words 4t .. 4t + 3 of each 1024-word pass
    int running = 0;
    for (int q0 = 0; q0 < nw; q0 += 1024) {
        int pc[4], sum = 0;
        for (int u = 0; u < 4; ++u) {
            int q = q0 + 4 * t + u;
            pc[u] = q < nw ? __popc(bits[q]) : 0;
            sum += pc[u];
        }
        int tot;
        int ex = block_scan_excl(sum, &tot) + running;
        for (int u = 0; u < 4; ++u) {
            int q = q0 + 4 * t + u;
            if (q < nw) wpre[q] = ex;
            ex += pc[u];
        }
        running += tot;
    }
    __syncthreads();
    if (running > HALO_MAX) {
        if (t == 0) c.h_count[b_] = -1;
        return;
    }
    const int gbase = c.ws[w] + lo;
    int* ids = c.h_ids + int64_t(b_) * HALO_MAX;
    for (int q = t; q < nw; q += 256) {
        unsigned m = bits[q];
        int r = wpre[q];
        while (m) {
            int b = __ffs(m) - 1;
            ids[r++] = gbase + 32 * q + b;
            m &= m - 1u;
        }
    }
    for (int k = 0; k < cnt; ++k) {
        int d = entry(k) - lo;
        unsigned below = bits[d >> 5] & ((1u << (d & 31)) - 1u);
        (c.h_slot + int64_t(k) * c.n)[i] = uint16_t(wpre[d >> 5] + __popc(below));
    }
    if (t == 0) c.h_count[b_] = running;
}

// Neighbour loop over a staged halo: the rows of halo indices (h_slot, padded like the rows) and the records in LDS
// (load(s), s < h_count); otherwise as for_rows_g: body(record, grad W_ij) in ascending k.
// SU entries per batch (8 measured slower than 4: more registers, fewer waves); the next batch's halo indices and
// cached scales are loaded before the current batch's arithmetic (the rows hold ROW_PAD >= 2 SU padding rows), so
// their memory latency overlaps it.
constexpr int SU = 4;
static_assert(ROW_PAD >= 2 * SU && ROW_PAD >= NB_U, "row padding");
template <class T, class LoadF, class BodyF>
__device__ __forceinline__ void for_slots_g(const Ctx& c, int i, f3 xi, int cnt, bool cached, LoadF load, BodyF body) {
    const KC K = make_kc(c.h, c.kind);
    int sl[SU];
    float sc[SU];
    auto fetch = [&](int k0, int (&s_)[SU], float (&g_)[SU]) {
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            s_[u] = int((c.h_slot + int64_t(k0 + u) * c.n)[i]);
            g_[u] = cached ? (c.pgrad + int64_t(k0 + u) * c.n)[i] : 0.0f;
        }
    };
    if (cnt > 0) fetch(0, sl, sc);
    for (int k0 = 0; k0 < cnt; k0 += SU) {
        T d[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) d[u] = load(k0 + u < cnt ? sl[u] : 0);
        int sl_n[SU];
        float sc_n[SU];
        fetch(k0 + SU, sl_n, sc_n);
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            if (k0 + u >= cnt) continue;
            f3 rv = xi - xyz(load_pos(d[u]));
            body(d[u], cached ? pair_grad(rv, sc[u]) : kernel_grad(rv, K));
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            sl[u] = sl_n[u];
            sc[u] = sc_n[u];
        }
    }
}

// Stage the chunk's halo into LDS: every index of the thread's share first, then every record, then the stores, so
// a block waits for two memory round trips rather than one pair per record.
constexpr int HALO_PER_THREAD = (HALO_MAX + 255) / 256;
template <class RecF, class PutF>
__device__ __forceinline__ void stage_halo(const Ctx& c, int b, int hc, RecF rec, PutF put) {
    if (hc <= 0) return;
    const int* ids = c.h_ids + int64_t(b) * HALO_MAX;
    int jj[HALO_PER_THREAD];
#pragma unroll
    for (int u = 0; u < HALO_PER_THREAD; ++u) {
        int t = int(threadIdx.x) + 256 * u;
        jj[u] = t < hc ? ids[t] : -1;
    }
    using R = decltype(rec(0));
    R r[HALO_PER_THREAD];
#pragma unroll
    for (int u = 0; u < HALO_PER_THREAD; ++u)
        if (jj[u] >= 0) r[u] = rec(jj[u]);
#pragma unroll
    for (int u = 0; u < HALO_PER_THREAD; ++u)
        if (jj[u] >= 0) put(int(threadIdx.x) + 256 * u, r[u]);
}

// pressure.py:144-193 for one particle (is_fluid), returning 1 for the chunk partial of `active` (solver.py:241).
// With FEAT the same neighbour pass also gathers the neighbours' velocities and writes the particle's learned
// warm-start features [16] (13 network inputs, then the rule's initial kappa of both solves), so the learned start
// costs no pass of its own: neighbour count / n_full, 100 (r - 1), 100 dt Dr/Dt, asinh(kappa_v,prev / ks),
// asinh(kappa_d,prev / ks), h grad r (3), sum psi, h |grad r|, alpha D, n_eff / n_full, h grad r . g_hat; ks = |g| h.
struct QVd {
    float4 q;
    f3 v;
};
template <bool WIDE, bool FEAT>
__device__ __forceinline__ void density_alpha_one(const Ctx& c, const KC& K, int i, int w, float dt) {
    float4 qi = c.q4[i];
    f3 xi = xyz(qi);
    float mi = qi.w;
    float inv_r0 = 1.0f / c.rho0[i];
    float rho = mi * kernel_w(0.0f, K);
    f3 G{0.0f, 0.0f, 0.0f}, vi{0.0f, 0.0f, 0.0f};
    float sq = 0.0f, rate = 0.0f;
    if (FEAT) vi = V(c, i);
    int cnt = c.count[i];
    const int base = c.ws[w];
    const unsigned sm = slot_mask(c, i);
    const WorldRecs q4 = WorldRecs::of(c.q4, base, c.n), qd = WorldRecs::of(reinterpret_cast<const f3*>(c.qd), base, c.n);
    auto loop = [&](auto kv) {  // the kernel variant chosen once for the loop (kernel_w_rv)
        constexpr int KV = decltype(kv)::value;
        for_rows<WIDE, QVd>(
            c, i, cnt, [&](unsigned e) { return QVd{q4.f4(e), FEAT ? qd.v3(e) : f3{0.0f, 0.0f, 0.0f}}; },
            [&](int k, QVd d) {
                f3 rv = xi - xyz(d.q);
                float mj = d.q.w;
                rho += mj * kernel_w_rv<KV>(rv, K);
                float sc = kernel_grad_scale_v<KV>(rv, K);
                if (c.pair_cache) c.pgrad[int64_t(k) * c.n + i] = sc;
                f3 g = pair_grad(rv, sc);
                G = G + (mj * inv_r0) * g;
                sq += (mi * mj * inv_r0 * inv_r0) * dot(g, g);
                if (FEAT) rate += (mj * inv_r0) * dot(vi - d.v, g);
            });
    };
    if (K.fast && K.kind == 0) loop(std::integral_constant<int, 0>{});
    else loop(std::integral_constant<int, 1>{});
    float psi = 0.0f, lam = 0.0f;
    int64_t b4 = 4 * int64_t(i);
    for (int b = 0; b < 4; ++b) {
        if (!((sm >> b) & 1u)) continue;
        psi += c.s_psi[b4 + b];
        f3 gb = ld3(c.s_gpsi, b4 + b);
        G = G + gb;
        if (FEAT) rate += dot(vi - ld3(c.s_vb, b4 + b), gb);
        for (int cc = b + 1; cc < 4; ++cc)
            if ((sm >> cc) & 1u) lam += fmx(-2.0f * dot(gb, ld3(c.s_gpsi, b4 + cc)), 0.0f);
    }
    float tot = dot(G, G) + sq + lam;
    lam += fmx(c.tot_min - tot, 0.0f);
    tot = fmx(tot, c.tot_min);
    float ratio = rho * inv_r0 + psi;
    float alpha = tot > 1.0e-12f ? 1.0f / tot : 0.0f;
    float neff = float(cnt) + psi * c.n_full;
    c.reg[i] = lam;
    c.ratio[i] = ratio;
    c.alpha[i] = alpha;
    c.neff[i] = neff;
    c.active[i] = 1.0f;
    if (FEAT) {
        float* F = c.feat + 16 * int64_t(i);
        f3 gw = ld3(c.gravity, w);
        float gl = length(gw);
        f3 gh = gl > 0.0f ? gw / gl : f3{0.0f, 0.0f, -1.0f};
        float ks = c.kscale;
        F[0] = float(cnt) / c.n_full;
        F[1] = 100.0f * (ratio - 1.0f);
        F[2] = 100.0f * dt * rate;
        F[3] = asinhf(c.kv[i] / dt / ks);
        F[4] = asinhf(c.kd[i] / (dt * dt) / ks);
        F[5] = c.h * G.x;
        F[6] = c.h * G.y;
        F[7] = c.h * G.z;
        F[8] = psi;
        F[9] = c.h * length(G);
        F[10] = alpha * c.lat_diag;
        F[11] = neff / c.n_full;
        F[12] = c.h * dot(G, gh);
        // not network inputs: the reference rule's initial kappa of each solve (pressure.py:254, 261, before the
        // lower bound), the baseline the network's output corrects
        F[13] = asinhf(c.warm_v * fmn(c.kv[i], 0.5f) / dt / ks);
        F[14] = asinhf(c.warm_d * fmn(c.kd[i], 2.5e-4f) / (dt * dt) / ks);
        F[15] = 0.0f;
    }
}
template <bool FEAT>
__device__ __forceinline__ void density_alpha_other(const Ctx& c, int i) {
    c.ratio[i] = 0.0f;
    c.alpha[i] = 0.0f;
    c.neff[i] = 0.0f;
    c.active[i] = 0.0f;
    c.reg[i] = 0.0f;
    if (FEAT)
        for (int k = 0; k < 16; ++k) c.feat[16 * int64_t(i) + k] = 0.0f;
}

// gated (the learned start's per-world gate): only the chunks of worlds whose learned start is off; the features pass
// takes the others
template <bool WIDE>
__global__ void __launch_bounds__(256) k_density_alpha(Ctx c, int gated) {
    const KC K = make_kc(c.h, c.kind);
    CHUNK_PROLOGUE
    if (gated && c.learn_on[c.ch_world[b_]]) return;
    double act_v = 0.0;
    if (valid) {
        if (!is_fluid(c, i)) {
            density_alpha_other<false>(c, i);
        } else {
            density_alpha_one<WIDE, false>(c, K, i, c.ch_world[b_], 0.0f);
            act_v = 1.0;
        }
    }
    chunk_reduce(c, b_, act_v, 0.0, valid ? act_v : -3.4e38);
}
// the same pass plus the learned warm start's features of the step (dt) in c.feat
template <bool WIDE>
__global__ void __launch_bounds__(256) k_density_features(Ctx c, float dt, int gated) {
    const KC K = make_kc(c.h, c.kind);
    CHUNK_PROLOGUE
    if (gated && !c.learn_on[c.ch_world[b_]]) return;
    double act_v = 0.0;
    if (valid) {
        if (!is_fluid(c, i)) {
            density_alpha_other<true>(c, i);
        } else {
            density_alpha_one<WIDE, true>(c, K, i, c.ch_world[b_], dt);
            act_v = 1.0;
        }
    }
    chunk_reduce(c, b_, act_v, 0.0, valid ? act_v : -3.4e38);
}

// ----------------------------------------------------------------------------------------------- solves
struct QV {
    float4 q;
    f3 v;
};
// w: the particle's world, block-uniform (the caller runs one chunk per block)
template <bool WIDE>
__device__ float density_rate(const Ctx& c, int i, int w) {
    f3 xi = xyz(c.q4[i]), vi = V(c, i);
    float inv_r0 = 1.0f / c.rho0[i];
    float rate = 0.0f;
    int cnt = c.count[i];
    const int base = c.ws[w];
    const unsigned sm = slot_mask(c, i);
    const WorldRecs q4 = WorldRecs::of(c.q4, base, c.n), qd = WorldRecs::of(reinterpret_cast<const f3*>(c.qd), base, c.n);
    for_rows_g<WIDE, QV>(c, i, xi, cnt, (c.pair_cache & DFSPH_PAIR_CACHE_OTHER) != 0,
                         [&](unsigned e) { return QV{q4.f4(e), qd.v3(e)}; },
                         [&](QV d, f3 g) { rate += (d.q.w * inv_r0) * dot(vi - d.v, g); });
    int64_t b4 = 4 * int64_t(i);
    for (int b = 0; b < 4; ++b)
        if ((sm >> b) & 1u) rate += dot(vi - ld3(c.s_vb, b4 + b), ld3(c.s_gpsi, b4 + b));
    return rate;
}

// pressure.py:214-261; block (0, 0) clears the unconverged-world count for begin_solve
template <bool WIDE>
__global__ void __launch_bounds__(256) k_solve_init(Ctx c, int mode, float dt, float warm, float warm_den,
                                                    float relief, const float* store, const float* learned) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *c.n_unconv = 0;
    CHUNK_PROLOGUE
    if (!valid) return;
    if (!is_fluid(c, i)) {
        c.src[i] = 0.0f;
        c.kappa[i] = 0.0f;
        put_vr(c, c.kr, VR_KAPPA, i, 0.0f);
        c.lower[i] = 0.0f;
        return;
    }
    // read before density_rate's neighbour loop, which orders its loads before what follows it
    const bool use_learned = learned && (!c.learn_gate || c.learn_on[c.ch_world[b_]]);
    const float store_i = store[i], learned_i = use_learned ? learned[i] : 0.0f, rho0_i = c.rho0[i];
    const float kd_i = mode == 0 && c.two_sided ? c.kd[i] : 0.0f, neff_i = mode == 0 ? c.neff[i] : 0.0f;
    const float ratio_i = mode == 0 ? 0.0f : c.ratio[i];
    float rate = density_rate<WIDE>(c, i, c.ch_world[b_]);
    float kap;
    if (mode == 0) {
        float kdv = 0.0f;
        if (c.two_sided) kdv = fmx(warm_den * fmn(kd_i, 2.5e-4f) / (dt * dt), 0.0f);
        float lb = -kdv;
        float rr = rate;
        if (kdv <= 0.0f) rr = fmx(rate, 0.0f);
        if (neff_i < c.min_nb) {
            rr = 0.0f;
            lb = 0.0f;
        }
        c.src[i] = rr;
        c.lower[i] = lb;
        kap = fmx(warm * fmn(store_i, 0.5f) / dt, lb);
        if (use_learned) kap = fmx(learned_i, lb);
    } else {
        float rs = ratio_i + dt * rate;
        if (ratio_i - 1.0f > relief) rs = 1.0f + relief + dt * rate;
        c.src[i] = rs;
        c.lower[i] = 0.0f;
        kap = warm * fmn(store_i, 2.5e-4f) / (dt * dt);
        if (use_learned) kap = fmx(learned_i, 0.0f);
    }
    c.kappa[i] = kap;
    put_vr(c, c.kr, VR_KAPPA, i, kap / rho0_i);
}

// pressure.py:355-367
__global__ void k_begin_solve(Ctx c) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W) return;
    c.iters[w] = 0;
    c.err_mean[w] = 0.0f;
    c.restart[w] = 0;
    if (c.n_active[w] > 0.0) {
        c.converged[w] = 0;
        atomicAdd(c.n_unconv, 1);
    } else {
        c.converged[w] = 1;
    }
}

// true when p[j] != 0 for a listed neighbour j of i (stops at the first one)
template <bool WIDE, class T>
__device__ __forceinline__ bool any_nbr_nonzero(const Ctx& c, int i, int cnt, int base, const T* p) {
    bool any = false;
    for (int k0 = 0; k0 < cnt && !any; k0 += NB_U) {
        int j[NB_U];
        nbr_batch<WIDE>(c, i, cnt, base, k0, j);
        T v[NB_U];
#pragma unroll
        for (int u = 0; u < NB_U; ++u) v[u] = p[j[u]];
#pragma unroll
        for (int u = 0; u < NB_U; ++u)
            if (k0 + u < cnt && v[u] != T(0)) any = true;
    }
    return any;
}

// pressure.py:264-277
struct QKR {
    float4 q;
    float kr;
};
// kapr = kap / rho0 per particle, written by the kernel that produced kap (the same division, gathered once)
// With c.packed (one mass per world) each neighbour is one 16 B gather {x, y, z, kap / rho0} of the vector's packed
// copy, the mass the world's; otherwise a position-and-mass gather plus the quotient's. The terms are the same floats.
template <bool WIDE>
__device__ f3 pressure_accel(const Ctx& c, const float* kap, const float* kapr, int slot, int i, int w) {
    f3 xi = xyz(c.q4[i]);
    const float kap_i = kap[i];
    float ki = kap_i / c.rho0[i];
    f3 a{0.0f, 0.0f, 0.0f};
    int cnt = c.count[i];
    const int base = c.ws[w];
    const bool cached = (c.pair_cache & DFSPH_PAIR_CACHE_ACCEL) != 0;
    const unsigned sm = slot_mask(c, i);
    if (c.packed) {
        const float mw = c.w_mass[w];
        const WorldRecs qk = WorldRecs::of(c.qk + int64_t(slot) * c.n, base, c.n);
        for_rows_g<WIDE, float4>(c, i, xi, cnt, cached, [&](unsigned e) { return qk.f4(e); },
                                 [&](float4 d, f3 g) { a = a - (mw * (ki + d.w)) * g; });
    } else {
        const WorldRecs q4 = WorldRecs::of(c.q4, base, c.n), kr = WorldRecs::of(kapr, base, c.n);
        for_rows_g<WIDE, QKR>(c, i, xi, cnt, cached, [&](unsigned e) { return QKR{q4.f4(e), kr.f1(e)}; },
                              [&](QKR d, f3 g) { a = a - (d.q.w * (ki + d.kr)) * g; });
    }
    int64_t b4 = 4 * int64_t(i);
    for (int b = 0; b < 4; ++b)
        if ((sm >> b) & 1u) a = a - kap_i * ld3(c.s_gpsi, b4 + b);
    return a;
}

// pressure.py:294-310
struct QA {
    float4 q, a;
};
struct QA2 {  // the packed record {x, y, z, acc.x}, {acc.y, acc.z}
    float4 q;
    float2 a;
};
template <bool WIDE>
__device__ float accel_divergence(const Ctx& c, int i, int w) {
    f3 xi = xyz(c.q4[i]), ai = acc_of(c, i);
    float inv_r0 = 1.0f / c.rho0[i];
    float ar = 0.0f;
    int cnt = c.count[i];
    const int base = c.ws[w];
    const bool cached = (c.pair_cache & DFSPH_PAIR_CACHE_APPLY) != 0;
    const unsigned sm = slot_mask(c, i);
    if (c.packed) {  // one 16 B and one 8 B gather per neighbour; its mass is the world's
        const float mw = c.w_mass[w];
        const WorldRecs qa = WorldRecs::of(c.qa, base, c.n), a2 = WorldRecs::of(c.a2, base, c.n);
        for_rows_g<WIDE, QA2>(c, i, xi, cnt, cached, [&](unsigned e) { return QA2{qa.f4(e), a2.f2(e)}; },
                              [&](QA2 d, f3 g) { ar += (mw * inv_r0) * dot(ai - f3{d.q.w, d.a.x, d.a.y}, g); });
    } else {
        const WorldRecs q4 = WorldRecs::of(c.q4, base, c.n), acc4 = WorldRecs::of(c.acc4, base, c.n);
        for_rows_g<WIDE, QA>(c, i, xi, cnt, cached, [&](unsigned e) { return QA{q4.f4(e), acc4.f4(e)}; },
                             [&](QA d, f3 g) { ar += (d.q.w * inv_r0) * dot(ai - xyz(d.a), g); });
    }
    int64_t b4 = 4 * int64_t(i);
    for (int b = 0; b < 4; ++b)
        if ((sm >> b) & 1u) ar += dot(ai, ld3(c.s_gpsi, b4 + b));
    return ar;
}

// Algorithmic (compulsory) bytes of one particle in a pressure-acceleration / operator-apply launch: each
// per-particle array the pass reads or writes, once -- a neighbour's position, kappa quotient or acceleration is
// counted in its owner's read -- plus the row entries (2 B, 3 B in 24-bit worlds) and 12 B per filled slot
// gradient. k_accel, liquid: q4 16, vec 4, vec / rho0 4, rho0 4, count 4, flags 4, phase 1, world 4, slot shapes 16,
// acc written 12; other particles: flags, phase, acc written (17). k_apply, liquid: q4 16, acc 12, rho0 4, count 4,
// flags 4, phase 1, world 4, reg 4, vec 4, out written 4, slot shapes 16 (+4 source with the bias, +4 out read
// when accumulating); other particles: flags, phase, out written (9). Padding and caches are not counted.
template <bool WIDE>
__device__ __forceinline__ unsigned pass_bytes(const Ctx& c, int i, unsigned liquid, unsigned other) {
    if (!is_fluid(c, i)) return other;
    int64_t b4 = 4 * int64_t(i);
    unsigned ns = 0;
    for (int b = 0; b < 4; ++b) ns += c.s_shape[b4 + b] >= 0 ? 1u : 0u;
    return liquid + (WIDE ? 3u : 2u) * unsigned(c.count[i]) + 12u * ns;
}

__device__ __forceinline__ bool block_skip(const Ctx& c, int w, int gated) {
    if (c.converged[w] != 0) return true;
    return gated != 0 && c.restart[w] == 0;
}

// acc4 from the packed acceleration records
__global__ void __launch_bounds__(256) k_unpack_acc(Ctx c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const float2 a = c.a2[i];
    c.acc4[i] = make_float4(c.qa[i].w, a.x, a.y, 0.0f);
}
// the packed copies {x, y, z, vec / rho0} of the three vectors from q4 and the quotient arrays, and the packed
// acceleration records from q4 and acc4
__global__ void __launch_bounds__(256) k_pack_sync(Ctx c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const float4 q = c.q4[i];
    const float* vr[3] = {c.kr, c.pdr, c.corrr};
    for (int s = 0; s < 3; ++s) c.qk[int64_t(s) * c.n + i] = make_float4(q.x, q.y, q.z, vr[s][i]);
    const float4 a = c.acc4[i];
    c.qa[i] = make_float4(q.x, q.y, q.z, a.x);
    c.a2[i] = make_float2(a.y, a.z);
}

// accel_kernel (pressure.py:280-291) / cg_accel_kernel (krylov.py:77-87)
template <bool WIDE>
__device__ __forceinline__ void accel_chunk(const Ctx& c, int b_, const float* vec, const float* vecr, int slot,
                                            int gated) {
    int i = c.ch_begin[b_] + threadIdx.x;
    bool valid = i < c.ch_end[b_];
    if (block_skip(c, c.ch_world[b_], gated) || !c.ch_live[b_]) return;
    unsigned nb = 0;  // algorithmic bytes of this particle (profiling)
    if (valid) {
        if (!is_fluid(c, i)) {
            put_acc(c, i, 0.0f, 0.0f, 0.0f);
            nb = 17u;
        } else {
            bool full = true;
            if (gated && c.sparse_corr) {
                // every term of pressure_accel is m_j (k_i + k_j) grad W or k_i grad psi: with vec_i = 0 (so k_i = 0)
                // and every gathered k_j = 0 each term is +-0 and the sum stays exactly +0, which is what is stored
                int cnt = c.count[i];
                full = vec[i] != 0.0f || any_nbr_nonzero<WIDE>(c, i, cnt, c.ws[c.world[i]], vecr);
                c.touched[i] = full ? 1 : 0;
                if (!full) {
                    put_acc(c, i, 0.0f, 0.0f, 0.0f);
                    // flags 4, phase 1, vec 4, vec / rho0 4, count 4, world 4, touched 1, acc written 12
                    nb = 34u + (WIDE ? 3u : 2u) * unsigned(cnt);
                }
            }
            if (full) {
                f3 a = pressure_accel<WIDE>(c, vec, vecr, slot, i, c.ch_world[b_]);
                put_acc(c, i, a.x, a.y, a.z);
                nb = pass_bytes<WIDE>(c, i, 69u + (gated && c.sparse_corr ? 1u : 0u), 17u);
            }
        }
    }
    if (c.prof_bytes) account_bytes(c, KID_ACCEL, b_, nb);
}
template <bool WIDE>
__global__ void __launch_bounds__(256) k_accel(Ctx c, const float* vec, const float* vecr, int slot, int gated) {
    accel_chunk<WIDE>(c, xcd_block(int(blockIdx.x), int(gridDim.x)), vec, vecr, slot, gated);
}

// cg_apply_kernel (krylov.py:90-109); with `reduce` the chunk partial of pair.x = vec_i (A vec)_i
template <bool WIDE>
__device__ __forceinline__ void apply_chunk(const Ctx& c, int b_, int mode, float dt, const float* vec, float* out,
                                            int gated, int bias, int accum, int reduce) {
    int i = c.ch_begin[b_] + threadIdx.x;
    bool valid = i < c.ch_end[b_];
    if (block_skip(c, c.ch_world[b_], gated)) return;
    if (!c.ch_live[b_]) {
        if (reduce) chunk_zero(c, b_);
        return;
    }
    unsigned nb = 0;  // algorithmic bytes of this particle (profiling)
    double px = 0.0;
    if (valid) {
        if (!is_fluid(c, i)) {
            out[i] = 0.0f;
            nb = 9u;
        } else {
            bool full = true;
            if (gated && c.sparse_corr) {
                // the gated pass accumulates A corr: where neither this particle nor a neighbour had its acceleration
                // evaluated (all of them exact +0), the divergence is +0, v = -0 and out += v leaves out unchanged
                int cnt = c.count[i];
                full = c.touched[i] != 0 || any_nbr_nonzero<WIDE>(c, i, cnt, c.ws[c.world[i]], c.touched);
                if (!full) nb = 14u + (WIDE ? 3u : 2u) * unsigned(cnt);  // flags, phase, touched, count, world
            }
            if (full) {
                float sc = mode == 0 ? dt : dt * dt;
                // read before accel_divergence's neighbour loop, which orders its loads before what follows it
                const float vec_i = vec[i], reg_i = c.reg[i];
                const float src_i = bias ? c.src[i] : 0.0f, out_i = accum ? out[i] : 0.0f;
                float v = -sc * (accel_divergence<WIDE>(c, i, c.ch_world[b_]) - reg_i * vec_i);
                px = double(vec_i * v);
                if (bias) v -= (mode == 0 ? src_i : src_i - 1.0f);
                if (accum) v += out_i;
                out[i] = v;
                nb = pass_bytes<WIDE>(c, i, 77u + (bias ? 4u : 0u) + (accum ? 4u : 0u) +
                                                (gated && c.sparse_corr ? 1u : 0u), 9u);
            }
        }
    }
    if (c.prof_bytes) account_bytes(c, KID_APPLY, b_, nb);
    if (reduce) chunk_reduce_x(c, b_, px, valid);
}
template <bool WIDE>
__global__ void __launch_bounds__(256) k_apply(Ctx c, int mode, float dt, const float* vec, float* out, int gated,
                                               int bias, int accum, int reduce) {
    apply_chunk<WIDE>(c, xcd_block(int(blockIdx.x), int(gridDim.x)), mode, dt, vec, out, gated, bias, accum, reduce);
}

// k_accel / k_apply with the chunk's halo staged in LDS (DFSPH_OPT_HALO). The records: PACKED {x, y, z, vec / rho0}
// from the vector's packed copy (the mass the world's), else the position-and-mass record and the quotient; for the
// apply the position-and-mass record and the acceleration. A chunk without a halo (h_count -1) gathers directly.
// The sparse-correction test of the gated pass (sparse_corr) still reads its neighbours' values directly.
template <bool WIDE, bool PACKED>
__device__ __forceinline__ void accel_h(const Ctx& c, const float* vec, const float* vecr, int slot, int gated) {
    __shared__ float4 sq[HALO_MAX];
    __shared__ float skr[PACKED ? 1 : HALO_MAX];
    CHUNK_PROLOGUE
    if (block_skip(c, c.ch_world[b_], gated) || !c.ch_live[b_]) return;
    const int hc = c.h_count[b_];
    if (PACKED) {
        const float4* qk = c.qk + int64_t(slot) * c.n;
        stage_halo(c, b_, hc, [&](int j) { return qk[j]; }, [&](int t, float4 r) { sq[t] = r; });
    } else {
        stage_halo(c, b_, hc, [&](int j) { return QKR{c.q4[j], vecr[j]}; }, [&](int t, QKR r) {
            sq[t] = r.q;
            skr[t] = r.kr;
        });
    }
    __syncthreads();
    unsigned nb = 0;
    if (valid) {
        if (!is_fluid(c, i)) {
            put_acc(c, i, 0.0f, 0.0f, 0.0f);
            nb = 17u;
        } else {
            bool full = true;
            if (gated && c.sparse_corr) {
                int cnt = c.count[i];
                full = vec[i] != 0.0f || any_nbr_nonzero<WIDE>(c, i, cnt, c.ws[c.world[i]], vecr);
                c.touched[i] = full ? 1 : 0;
                if (!full) {
                    put_acc(c, i, 0.0f, 0.0f, 0.0f);
                    nb = 34u + (WIDE ? 3u : 2u) * unsigned(cnt);
                }
            }
            if (full) {
                f3 a;
                if (hc < 0) {
                    a = pressure_accel<WIDE>(c, vec, vecr, slot, i, c.ch_world[b_]);
                } else {  // pressure_accel over the staged records
                    f3 xi = xyz(c.q4[i]);
                    float ki = vec[i] / c.rho0[i];
                    a = f3{0.0f, 0.0f, 0.0f};
                    const bool cached = (c.pair_cache & DFSPH_PAIR_CACHE_ACCEL) != 0;
                    if (PACKED) {
                        const float mw = c.w_mass[c.world[i]];
                        for_slots_g<float4>(c, i, xi, c.count[i], cached, [&](int s) { return sq[s]; },
                                            [&](float4 d, f3 g) { a = a - (mw * (ki + d.w)) * g; });
                    } else {
                        for_slots_g<QKR>(c, i, xi, c.count[i], cached, [&](int s) { return QKR{sq[s], skr[s]}; },
                                         [&](QKR d, f3 g) { a = a - (d.q.w * (ki + d.kr)) * g; });
                    }
                    int64_t b4 = 4 * int64_t(i);
                    for (int b = 0; b < 4; ++b)
                        if (c.s_shape[b4 + b] >= 0) a = a - vec[i] * ld3(c.s_gpsi, b4 + b);
                }
                put_acc(c, i, a.x, a.y, a.z);
                nb = pass_bytes<WIDE>(c, i, 69u + (gated && c.sparse_corr ? 1u : 0u), 17u);
            }
        }
    }
    if (c.prof_bytes) account_bytes(c, KID_ACCEL, b_, nb);
}
template <bool WIDE>
__global__ void __launch_bounds__(256) k_accel_hp(Ctx c, const float* vec, const float* vecr, int slot, int gated) {
    accel_h<WIDE, true>(c, vec, vecr, slot, gated);
}
template <bool WIDE>
__global__ void __launch_bounds__(256) k_accel_hn(Ctx c, const float* vec, const float* vecr, int slot, int gated) {
    accel_h<WIDE, false>(c, vec, vecr, slot, gated);
}

// PACKED (one particle mass per world, the packed acceleration records' condition): 24 B records {x, y, z, a_x} and
// {a_y, a_z}, the neighbour's mass the world's; else {x, y, z, m} and the acceleration
template <bool WIDE, bool PACKED>
__device__ __forceinline__ void apply_h(const Ctx& c, int mode, float dt, const float* vec, float* out, int gated,
                                        int bias, int accum, int reduce) {
    __shared__ float4 sq[HALO_MAX];
    __shared__ float4 sa[PACKED ? 1 : HALO_MAX];
    __shared__ float2 sa2[PACKED ? HALO_MAX : 1];
    CHUNK_PROLOGUE
    if (block_skip(c, c.ch_world[b_], gated)) return;
    if (!c.ch_live[b_]) {
        if (reduce) chunk_zero(c, b_);
        return;
    }
    const int hc = c.h_count[b_];
    if (PACKED) {
        stage_halo(c, b_, hc, [&](int j) { return QA2{c.qa[j], c.a2[j]}; }, [&](int t, QA2 r) {
            sq[t] = r.q;
            sa2[t] = r.a;
        });
    } else {
        stage_halo(c, b_, hc, [&](int j) { return QA{c.q4[j], c.acc4[j]}; }, [&](int t, QA r) {
            sq[t] = r.q;
            sa[t] = r.a;
        });
    }
    __syncthreads();
    unsigned nb = 0;
    double px = 0.0;
    if (valid) {
        if (!is_fluid(c, i)) {
            out[i] = 0.0f;
            nb = 9u;
        } else {
            bool full = true;
            if (gated && c.sparse_corr) {
                int cnt = c.count[i];
                full = c.touched[i] != 0 || any_nbr_nonzero<WIDE>(c, i, cnt, c.ws[c.world[i]], c.touched);
                if (!full) nb = 14u + (WIDE ? 3u : 2u) * unsigned(cnt);
            }
            if (full) {
                float div;
                if (hc < 0) {
                    div = accel_divergence<WIDE>(c, i, c.ch_world[b_]);
                } else {  // accel_divergence over the staged records
                    f3 xi = xyz(c.q4[i]), ai = acc_of(c, i);
                    float inv_r0 = 1.0f / c.rho0[i];
                    float ar = 0.0f;
                    const bool cached = (c.pair_cache & DFSPH_PAIR_CACHE_APPLY) != 0;
                    if (PACKED) {
                        const float mw = c.w_mass[c.world[i]];
                        for_slots_g<QA2>(c, i, xi, c.count[i], cached, [&](int s) { return QA2{sq[s], sa2[s]}; },
                                         [&](QA2 d, f3 g) {
                                             ar += (mw * inv_r0) * dot(ai - f3{d.q.w, d.a.x, d.a.y}, g);
                                         });
                    } else {
                        for_slots_g<QA>(c, i, xi, c.count[i], cached, [&](int s) { return QA{sq[s], sa[s]}; },
                                        [&](QA d, f3 g) { ar += (d.q.w * inv_r0) * dot(ai - xyz(d.a), g); });
                    }
                    int64_t b4 = 4 * int64_t(i);
                    for (int b = 0; b < 4; ++b)
                        if (c.s_shape[b4 + b] >= 0) ar += dot(ai, ld3(c.s_gpsi, b4 + b));
                    div = ar;
                }
                float sc = mode == 0 ? dt : dt * dt;
                float v = -sc * (div - c.reg[i] * vec[i]);
                px = double(vec[i] * v);
                if (bias) v -= (mode == 0 ? c.src[i] : c.src[i] - 1.0f);
                if (accum) v += out[i];
                out[i] = v;
                nb = pass_bytes<WIDE>(c, i, 77u + (bias ? 4u : 0u) + (accum ? 4u : 0u) +
                                                (gated && c.sparse_corr ? 1u : 0u), 9u);
            }
        }
    }
    if (c.prof_bytes) account_bytes(c, KID_APPLY, b_, nb);
    if (reduce) chunk_reduce_x(c, b_, px, valid);
}
template <bool WIDE>
__global__ void __launch_bounds__(256) k_apply_hp(Ctx c, int mode, float dt, const float* vec, float* out, int gated,
                                                  int bias, int accum, int reduce) {
    apply_h<WIDE, true>(c, mode, dt, vec, out, gated, bias, accum, reduce);
}
template <bool WIDE>
__global__ void __launch_bounds__(256) k_apply_hn(Ctx c, int mode, float dt, const float* vec, float* out, int gated,
                                                  int bias, int accum, int reduce) {
    apply_h<WIDE, false>(c, mode, dt, vec, out, gated, bias, accum, reduce);
}

// cg_init_kernel (krylov.py:112-132)
__global__ void __launch_bounds__(256) k_cg_init(Ctx c, int mode, float dt) {
    CHUNK_PROLOGUE
    if (c.converged[c.ch_world[b_]] != 0) return;
    if (!c.ch_live[b_]) {
        chunk_zero(c, b_);
        return;
    }
    double px = 0.0;
    if (valid) {
        if (!is_fluid(c, i)) {
            c.act[i] = 0;
            c.z[i] = 0.0f;
            c.pd[i] = 0.0f;
            put_vr(c, c.pdr, VR_DIR, i, 0.0f);
        } else {
            float sc = mode == 0 ? dt : dt * dt;
            int a = 0;
            float zi = 0.0f;
            float wi = c.wv[i];
            if (c.kappa[i] > c.lower[i] || wi < 0.0f) {
                a = 1;
                zi = -wi * c.alpha[i] / sc;
            }
            c.act[i] = a;
            c.z[i] = zi;
            c.pd[i] = zi;
            put_vr(c, c.pdr, VR_DIR, i, zi / c.rho0[i]);
            px = double(-wi * zi);
        }
    }
    chunk_reduce_x(c, b_, px, valid);
}

// cg_step_kernel (krylov.py:153-177)
__device__ __forceinline__ void cg_step_chunk(const Ctx& c, int b_) {
    int i = c.ch_begin[b_] + threadIdx.x;
    bool valid = i < c.ch_end[b_];
    int w = c.ch_world[b_];
    if (c.converged[w] != 0 || !valid || !c.ch_live[b_]) return;
    // corr and its quotients are written together, so where corr is +0 and stays +0 they hold +0 already (the
    // packed record's w is also set to +0 at each list build) and are not written again
    const bool corr_zero = __float_as_uint(c.corr[i]) == 0u;
    if (!is_fluid(c, i)) {
        if (!corr_zero) {
            c.corr[i] = 0.0f;
            put_vr(c, c.corrr, VR_CORR, i, 0.0f);
        }
        return;
    }
    double pap = c.sums[2 * w];
    double a64 = 0.0;
    if (pap > 1.0e-30) a64 = c.rz[w] / pap;
    float a = float(a64);
    float cc = 0.0f;
    if (c.act[i] != 0) {
        float k = c.kappa[i] + a * c.pd[i];
        if (k < c.lower[i]) {
            cc = c.lower[i] - k;
            k = c.lower[i];
            c.act[i] = 0;
            c.restart[w] = 1;  // a flag: only restart != 0 is ever tested (krylov.py:70-74, 231)
        }
        c.kappa[i] = k;
        put_vr(c, c.kr, VR_KAPPA, i, k / c.rho0[i]);
    }
    if (__float_as_uint(cc) != 0u || !corr_zero) {
        c.corr[i] = cc;
        put_vr(c, c.corrr, VR_CORR, i, cc / c.rho0[i]);
    }
    c.wv[i] = c.wv[i] + a * c.ap[i];
}
__global__ void __launch_bounds__(256) k_cg_step(Ctx c) { cg_step_chunk(c, xcd_block(int(blockIdx.x), int(gridDim.x))); }

// cg_finish_kernel (krylov.py:180-212) + the chunk partial of (r.z, error, max error)
__device__ __forceinline__ void cg_finish_chunk(const Ctx& c, int b_, int mode, float dt) {
    int i = c.ch_begin[b_] + threadIdx.x;
    bool valid = i < c.ch_end[b_];
    int w = c.ch_world[b_];
    if (c.converged[w] != 0) return;
    if (!c.ch_live[b_]) {  // the error of a particle that is not liquid is 0
        if (valid) c.err[i] = 0.0f;
        chunk_zero(c, b_);
        return;
    }
    double px = 0.0, py = 0.0;
    if (valid) {
        if (!is_fluid(c, i)) {
            c.err[i] = 0.0f;
        } else {
            float sc = mode == 0 ? dt : dt * dt;
            float wvv = c.wv[i];
            if (c.act[i] == 0 && wvv < 0.0f) {
                c.act[i] = 1;
                c.restart[w] = 1;
            }
            float zi = 0.0f;
            if (c.act[i] != 0) zi = -wvv * c.alpha[i] / sc;
            c.z[i] = zi;
            float e = fmx(-wvv, 0.0f);
            if (c.kappa[i] > c.lower[i]) e = fabsf(wvv);
            if (mode == 0) {
                if (c.neff[i] < c.min_nb) e = 0.0f;
                e *= dt;
            }
            c.err[i] = e;
            px = double(-wvv * zi);
            py = double(e);
        }
    }
    chunk_reduce(c, b_, px, py, valid ? py : -3.4e38);
}
__global__ void __launch_bounds__(256) k_cg_finish(Ctx c, int mode, float dt) {
    cg_finish_chunk(c, xcd_block(int(blockIdx.x), int(gridDim.x)), mode, dt);
}

// cg_direction_kernel (krylov.py:243-255)
__device__ __forceinline__ void cg_direction_chunk(const Ctx& c, int b_) {
    int i = c.ch_begin[b_] + threadIdx.x;
    bool valid = i < c.ch_end[b_];
    int w = c.ch_world[b_];
    if (c.converged[w] != 0 || !valid || !c.ch_live[b_] || !is_fluid(c, i)) return;
    const float p_old = c.pd[i];
    float p = 0.0f;
    if (c.act[i] != 0) p = c.z[i] + float(c.stepb[w]) * p_old;
    // p and its quotients are written together, so a stored +0 that stays +0 is not written again (the packed
    // record's w is also set to +0 at each list build)
    if (__float_as_uint(p) == 0u && __float_as_uint(p_old) == 0u) return;
    c.pd[i] = p;
    put_vr(c, c.pdr, VR_DIR, i, p / c.rho0[i]);
}
// flag_out (pinned host memory, may be null): block 0 stores the unconverged-world count that k_world's stop test
// just produced, so the host learns it from this batch's completion event without a separate copy
__global__ void __launch_bounds__(256) k_cg_direction(Ctx c, int* flag_out) {
    if (flag_out && blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(flag_out, *c.n_unconv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    cg_direction_chunk(c, xcd_block(int(blockIdx.x), int(gridDim.x)));
}

// ----------------------------------------------------------------------------------------------- device solve loop
// The projected-CG iterations of one solve (krylov.py:298-318's capture_while) as one cooperative launch: every
// block of the grid is resident, the phases of an iteration (the same per-chunk and per-world bodies the separate
// kernels run, in the same order) are separated by grid barriers, and the loop ends when the stop test of the last
// iteration left no unconverged world. No host round trip: the host enqueues the head, this kernel and the tail.

constexpr long long LOOP_WAIT_TICKS = 200000000LL;  // 2 s of the 100 MHz real-time counter: no phase takes that long

// Grid barrier k (k = 1, 2, ... within one launch), hierarchical over the XCDs (block b runs on XCD b % 8): thread 0
// of each block adds an arrival to its XCD's counter; the XCD's last arrival adds one to the global counter; the last
// of those stores k into every XCD's release flag; each block waits for its XCD's flag. The counters and flags sit
// on cache lines of their own (LOOP_LINE ints apart) and hold monotonic values, so nothing is reset between
// barriers. Read-modify-writes are acquire-release and the flags release / acquire at agent scope, so every write of
// the phase before the barrier is visible to every block after it. The wait is bounded: past LOOP_WAIT_TICKS, or
// when another block already gave up, the block sets the fault flag and the call returns false, and the kernel ends.
// `*watch` (may be null) is read by thread 0 after the barrier: every block sees the same value when no phase writes
// it between two barriers.
constexpr int LOOP_LINE = 32;                        // 128 B
constexpr int LOOP_WORDS = (2 * N_XCD + 1) * LOOP_LINE;  // XCD counters, the global counter, XCD release flags
__device__ __forceinline__ bool grid_barrier(const Ctx& c, unsigned k, const int* watch, int* seen) {
    __shared__ int s_ok, s_val;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned G = gridDim.x, x = blockIdx.x % N_XCD, per = G / N_XCD;
        unsigned* xcd_count = c.loop_bar + x * LOOP_LINE;
        unsigned* all_count = c.loop_bar + N_XCD * LOOP_LINE;
        unsigned* release = c.loop_bar + (N_XCD + 1 + x) * LOOP_LINE;
        unsigned old = __hip_atomic_fetch_add(xcd_count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1u == per * k) {  // this XCD's last arrival
            unsigned o2 = __hip_atomic_fetch_add(all_count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (o2 + 1u == N_XCD * k)  // the grid's last arrival
                for (int y = 0; y < N_XCD; ++y)
                    __hip_atomic_store(c.loop_bar + (N_XCD + 1 + y) * LOOP_LINE, k, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_AGENT);
        }
        const long long t0 = wall_clock64();
        int ok = 1;
        // relaxed polling (an acquire per poll would invalidate the caches on every read), one acquire fence after
        while (__hip_atomic_load(release, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k) {
            if (__hip_atomic_load(c.loop_fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                wall_clock64() - t0 > LOOP_WAIT_TICKS) {
                __hip_atomic_store(c.loop_fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(4);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
        s_val = watch ? __hip_atomic_load(watch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    }
    __syncthreads();
    if (seen) *seen = s_val;
    return s_ok != 0;
}

// body(chunk) over this block's share of the chunks: the blocks of XCD x (block b runs on XCD b % 8) take the
// contiguous x-th eighth of the chunk table, as xcd_block arranges it for the separate launches
template <class F>
__device__ __forceinline__ void loop_chunks(const Ctx& c, F body) {
    const int G = int(gridDim.x), x = int(blockIdx.x) % N_XCD, per = G / N_XCD;
    const int lo = int((int64_t(c.n_chunks) * x) / N_XCD), hi = int((int64_t(c.n_chunks) * (x + 1)) / N_XCD);
    for (int b = lo + int(blockIdx.x) / N_XCD; b < hi; b += per) body(b);
}

template <bool WIDE>
__global__ void __launch_bounds__(256) k_cg_loop(Ctx c, int mode, float dt, StopSpec sp) {
    const unsigned G = gridDim.x;
    unsigned target = 0;  // the barrier count
    int unconv = 0;
    if (!grid_barrier(c, ++target, c.n_unconv, &unconv)) return;
    for (int it = 0; unconv != 0; ++it) {
        if (it > 100000) {  // the stop test caps every world at max_iterations: never reached
            if (threadIdx.x == 0) __hip_atomic_store(c.loop_fault, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        loop_chunks(c, [&](int b) { accel_chunk<WIDE>(c, b, c.pd, c.pdr, VR_DIR, 0); });
        if (!grid_barrier(c, ++target, nullptr, nullptr)) return;
        loop_chunks(c, [&](int b) { apply_chunk<WIDE>(c, b, mode, dt, c.pd, c.ap, 0, 0, 0, 1); });
        if (!grid_barrier(c, ++target, nullptr, nullptr)) return;
        for (int w = int(blockIdx.x); w < c.W; w += int(G)) world_body(c, w, W_CG_PAP, sp);
        if (!grid_barrier(c, ++target, nullptr, nullptr)) return;
        loop_chunks(c, [&](int b) { cg_step_chunk(c, b); });
        if (!grid_barrier(c, ++target, nullptr, nullptr)) return;
        loop_chunks(c, [&](int b) { accel_chunk<WIDE>(c, b, c.corr, c.corrr, VR_CORR, 1); });
        if (!grid_barrier(c, ++target, nullptr, nullptr)) return;
        loop_chunks(c, [&](int b) { apply_chunk<WIDE>(c, b, mode, dt, c.corr, c.wv, 1, 0, 1, 0); });
        if (!grid_barrier(c, ++target, nullptr, nullptr)) return;
        loop_chunks(c, [&](int b) { cg_finish_chunk(c, b, mode, dt); });
        if (!grid_barrier(c, ++target, nullptr, nullptr)) return;
        for (int w = int(blockIdx.x); w < c.W; w += int(G)) world_body(c, w, W_CG_BETA, sp);
        if (!grid_barrier(c, ++target, nullptr, nullptr)) return;
        loop_chunks(c, [&](int b) { cg_direction_chunk(c, b); });
        if (!grid_barrier(c, ++target, c.n_unconv, &unconv)) return;
    }
}

// jacobi_kernel (pressure.py:313-352) + the chunk partial of (error, -, max error); block (0,0) clears the
// unconverged count the world epilogue recounts
template <bool WIDE>
__global__ void __launch_bounds__(256) k_jacobi(Ctx c, int mode, float dt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *c.n_unconv = 0;
    CHUNK_PROLOGUE
    int w = c.ch_world[b_];
    if (c.converged[w] != 0) return;
    double e64 = 0.0;
    if (valid) {
        if (!is_fluid(c, i)) {
            c.err[i] = 0.0f;
        } else {
            // read before accel_divergence's neighbour loop, which orders its loads before what follows it
            const float kap_i = c.kappa[i], reg_i = c.reg[i], src_i = c.src[i], lb = c.lower[i];
            const float alpha_i = c.alpha[i], rho0_i = c.rho0[i], neff_i = c.neff[i];
            float ar = accel_divergence<WIDE>(c, i, w) - reg_i * kap_i;
            float s, scale;
            if (mode == 0) {
                ar *= dt;
                s = -src_i;
                scale = 1.0f / dt;
            } else {
                ar *= dt * dt;
                s = 1.0f - src_i;
                scale = 1.0f / (dt * dt);
            }
            float rr = s - ar;
            float k_new = fmx(kap_i - c.omega * rr * alpha_i * scale, lb);
            c.kappa[i] = k_new;
            put_vr(c, c.kr, VR_KAPPA, i, k_new / rho0_i);
            float e = fmx(-rr, 0.0f);
            if (k_new > lb) e = fabsf(rr);
            if (mode == 0) {
                if (neff_i < c.min_nb) e = 0.0f;
                e *= dt;
            }
            c.err[i] = e;
            e64 = double(e);
        }
    }
    chunk_reduce(c, b_, e64, 0.0, valid ? e64 : -3.4e38);
}

// contact_limit_kernel + apply_pressure_kernel + lone_friction_kernel (pressure.py:393-480) for one particle: a
// lone particle (count 0) is nobody's neighbour, so limiting its kappa before the neighbours' pressure
// accelerations read kappa changes nothing for them.
template <bool WIDE>
__global__ void __launch_bounds__(256) k_post_solve(Ctx c, int mode, float dt, float* store) {
    CHUNK_PROLOGUE
    if (!valid || !is_fluid(c, i)) return;
    int64_t b4 = 4 * int64_t(i);
    bool lone = c.count[i] == 0;
    // read before pressure_accel's neighbour loop, which orders its loads before what follows it
    const f3 v0 = V(c, i);
    const float m_i = c.mass[i];
    const unsigned sm = slot_mask(c, i);
    float kap = c.kappa[i];
    if (lone && kap > 0.0f) {  // contact_limit_kernel
        f3 G{0.0f, 0.0f, 0.0f}, vb{0.0f, 0.0f, 0.0f};
        float wsum = 0.0f;
        for (int b = 0; b < 4; ++b)
            if ((sm >> b) & 1u) {
                G = G + ld3(c.s_gpsi, b4 + b);
                vb = vb + c.s_psi[b4 + b] * ld3(c.s_vb, b4 + b);
                wsum += c.s_psi[b4 + b];
            }
        float g = length(G);
        if (g > 0.0f) {
            if (wsum > 0.0f) vb = vb / wsum;
            float u = dot(v0 - vb, (-G) / g);
            kap = fmn(kap, fmx(c.v_sep - u, 0.0f) / (dt * g));
            c.kappa[i] = kap;
        }
    }
    // apply_pressure_kernel
    f3 a = pressure_accel<WIDE>(c, c.kappa, c.kr, VR_KAPPA, i, c.ch_world[b_]);
    f3 v = v0 + dt * a;
    float mk = m_i * kap * dt;
    for (int b = 0; b < 4; ++b)
        if ((sm >> b) & 1u) st3(c.s_imp, b4 + b, ld3(c.s_imp, b4 + b) + mk * ld3(c.s_gpsi, b4 + b));
    store[i] = mode == 0 ? kap * dt : kap * dt * dt;
    // lone_friction_kernel
    if (c.friction && lone && kap > 0.0f) {
        f3 G{0.0f, 0.0f, 0.0f}, vb{0.0f, 0.0f, 0.0f};
        float wsum = 0.0f, pbest = 0.0f;
        int best = -1;
        for (int b = 0; b < 4; ++b) {
            if ((sm >> b) & 1u) {
                float ps = c.s_psi[b4 + b];
                G = G + ld3(c.s_gpsi, b4 + b);
                vb = vb + ps * ld3(c.s_vb, b4 + b);
                wsum += ps;
                if (ps > pbest) {
                    pbest = ps;
                    best = b;
                }
            }
        }
        float g = length(G);
        if (g > 0.0f && wsum > 0.0f && best >= 0) {
            int s = c.s_shape[b4 + best];
            if (s < c.S) {
                float mu = c.sh_mu[s];
                if (mu > 0.0f) {
                    f3 nn = (-G) / g;
                    f3 u = v - vb / wsum;
                    f3 vt = u - dot(u, nn) * nn;
                    float lt = length(vt);
                    if (lt > 0.0f) {
                        f3 dv = fmn(1.0f, mu * dt * kap * g / lt) * vt;
                        v = v - dv;
                        st3(c.s_imp, b4 + best, ld3(c.s_imp, b4 + best) + m_i * dv);
                    }
                }
            }
        }
    }
    st3(c.qd, i, v);
}

// stats.record_solve / record_max_density_error / record_density_tolerance (runtime.py:112-122)
__global__ void k_record(Ctx c, int mode, float tol) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W) return;
    if (mode == 0) {
        c.it_div[w] = c.iters[w];
        c.st_div_err[w] = c.err_mean[w];
    } else {
        c.it_den[w] = c.iters[w];
        c.st_den_err[w] = c.err_mean[w];
        c.st_den_err_max[w] = c.n_active[w] > 0.0 ? c.err_max[w] : 0.0f;
        c.st_den_tol[w] = tol * c.tol_scale[w];
    }
}

// ----------------------------------------------------------------------------------------------- resting-jitter rule
// flow_terms_kernel (solve_driver.py:49-60): which 0 -> (v_z, 1), 1 -> (v_x, v_y)
__global__ void __launch_bounds__(256) k_flow_terms(Ctx c, int which) {
    CHUNK_PROLOGUE
    double x = 0.0, y = 0.0;
    if (valid && is_fluid(c, i)) {
        f3 v = V(c, i);
        if (which == 0) {
            x = v.z;
            y = 1.0;
        } else {
            x = v.x;
            y = v.y;
        }
    }
    chunk_reduce(c, b_, x, y, valid ? y : -3.4e38);
}
// head_terms_kernel (solve_driver.py:34-46)
__global__ void __launch_bounds__(256) k_head_terms(Ctx c) {
    CHUNK_PROLOGUE
    double x = 0.0, y = 0.0;
    if (valid && is_fluid(c, i)) {
        float k = fmx(c.kd[i], 0.0f);
        x = k;
        y = double(k * k);
    }
    chunk_reduce(c, b_, x, y, valid ? y : -3.4e38);
}
// tol_scale_kernel (solve_driver.py:63-82)
__global__ void k_tol_scale(Ctx c, float dt, float speed, float ratio, float tol, float tol_min) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W) return;
    float s = 1.0f;
    float g = length(ld3(c.gravity, w));
    float s1 = float(c.head[2 * w]), s2 = float(c.head[2 * w + 1]);
    float v = speed;
    double nn = c.flow_z[2 * w + 1];
    if (ratio > 0.0f && nn > 0.0) {
        f3 u{float(c.flow_xy[2 * w] / nn), float(c.flow_xy[2 * w + 1] / nn), float(c.flow_z[2 * w] / nn)};
        v = fmx(v, ratio * length(u));
    }
    if (g > 0.0f && s1 > 0.0f && s2 > 0.0f) {
        float depth = 1.5f * s2 / (s1 * dt * dt * g);
        s = fcl(v * dt / (depth * tol), tol_min / tol, 1.0f);
    }
    c.tol_scale[w] = s;
}

// ----------------------------------------------------------------------------------------------- non-pressure
// nonpressure_kernel (integrate.py:150-178)
struct XS {
    float4 q, vr;  // position and mass; velocity and density ratio * rho0 (k_pack_vrho)
};
// the XSPH records (v, ratio rho0): the product is the one the XSPH term forms for the neighbour, rounded once here
__global__ void __launch_bounds__(256) k_pack_vrho(Ctx c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const f3 v = V(c, i);
    c.vrho[i] = make_float4(v.x, v.y, v.z, c.ratio[i] * c.rho0[i]);
}
template <bool WIDE>
__global__ void __launch_bounds__(256) k_nonpressure(Ctx c, float dt) {
    const KC K = make_kc(c.h, c.kind);
    CHUNK_PROLOGUE
    if (!valid) return;
    if (!is_fluid(c, i)) return;  // the velocity of a particle that is not liquid is left as it is
    f3 vi = V(c, i);
    const int w = c.ch_world[b_];
    f3 a = ld3(c.gravity, w);
    if (c.f) a = a + ld3(c.f, i) / c.mass[i];
    f3 v = vi + dt * a;
    if (c.xsph > 0.0f) {
        f3 xi = xyz(c.q4[i]);
        f3 cr{0.0f, 0.0f, 0.0f};
        float rho_i = c.ratio[i] * c.rho0[i];
        int cnt = c.count[i];
        const int base = c.ws[w];
        const WorldRecs q4 = WorldRecs::of(c.q4, base, c.n), vr = WorldRecs::of(c.vrho, base, c.n);
        auto loop = [&](auto kv) {  // the kernel variant chosen once for the loop (kernel_w_rv)
            constexpr int KV = decltype(kv)::value;
            for_rows<WIDE, XS>(
                c, i, cnt, [&](unsigned e) { return XS{q4.f4(e), vr.f4(e)}; },
                [&](int, XS d) {
                    float rho_bar = 0.5f * (rho_i + d.vr.w);
                    if (rho_bar > 0.0f) {
                        float wk = kernel_w_rv<KV>(xi - xyz(d.q), K);
                        cr = cr + (d.q.w / rho_bar * wk) * (xyz(d.vr) - vi);
                    }
                });
        };
        if (K.fast && K.kind == 0) loop(std::integral_constant<int, 0>{});
        else loop(std::integral_constant<int, 1>{});
        v = v + c.xsph * cr;
    }
    // in place: the XSPH sums gather the velocities from the packed copy k_pack_vrho made before this pass
    st3(c.qd, i, v);
    st3(c.vtmp, i, v);
}

// ----------------------------------------------------------------------------------------------- advect / project
__device__ float path_sdf(const Ctx& c, int s, f3 l0, f3 l1, float t) {
    float phi;
    f3 g;
    shape_sdf(c.sh_geo[s], ld3(c.sh_scale, s), l0 + t * (l1 - l0), phi, g);
    return phi;
}
// box_span (integrate.py:66-81)
__device__ void box_span(f3 half, f3 l0, f3 l1, float& lo, float& hi) {
    lo = -1.0e30f;
    hi = 1.0e30f;
    f3 d = l1 - l0;
    float hv[3] = {half.x, half.y, half.z}, a0[3] = {l0.x, l0.y, l0.z}, dv[3] = {d.x, d.y, d.z};
    for (int a = 0; a < 3; ++a) {
        if (fabsf(dv[a]) > 0.0f) {
            float ta = (-hv[a] - a0[a]) / dv[a];
            float tb = (hv[a] - a0[a]) / dv[a];
            lo = fmx(lo, fmn(ta, tb));
            hi = fmn(hi, fmx(ta, tb));
        } else if (fabsf(a0[a]) > hv[a]) {
            lo = 2.0e30f;
        }
    }
}
// path_entry (integrate.py:84-114)
__device__ float path_entry(const Ctx& c, int s, f3 l0, f3 l1, int samples) {
    if (path_sdf(c, s, l0, l1, 0.0f) < 0.0f) return 2.0f;
    if (c.sh_geo[s] == GEO_BOX) {
        float lo, hi;
        box_span(ld3(c.sh_scale, s), l0, l1, lo, hi);
        if (lo <= hi && lo <= 1.0f && hi >= 0.0f) return fmx(lo, 0.0f);
        return 2.0f;
    }
    float t_out = 2.0f, t_lo = 0.0f, t_hi = 2.0f;
    for (int k = 1; k <= samples; ++k) {
        float t = float(k) / float(samples);
        if (t_hi > 1.0f) {
            if (path_sdf(c, s, l0, l1, t) < 0.0f) t_hi = t;
            else t_lo = t;
        }
    }
    if (t_hi <= 1.0f) {
        for (int b = 0; b < 8; ++b) {
            float tm = 0.5f * (t_lo + t_hi);
            if (path_sdf(c, s, l0, l1, tm) < 0.0f) t_hi = tm;
            else t_lo = tm;
        }
        t_out = t_lo;
    }
    return t_out;
}
// sweep_entry (integrate.py:125-147), static shapes
__device__ int sweep_entry(const Ctx& c, int w, f3 x0, f3 x1, f3& l_hit) {
    float len = length(x1 - x0);
    float reach = len + c.r;
    int samples = min(max(int(ceilf(len / c.r)) + 1, 8), 64);
    float best = 2.0f;
    int hit = -1;
    l_hit = f3{0.0f, 0.0f, 0.0f};
    for (int e = c.sh_start[w]; e < c.sh_start[w + 1]; ++e) {
        int s = c.sh_ids[e];
        Xf X = load_xf(c.sh_x + 7 * s);
        bool near_ = c.sh_geo[s] == GEO_PLANE || length(x0 - X.p) <= c.sh_radius[s] + reach;
        if (!near_) continue;
        f3 l0 = xf_inv_point(X, x0), l1 = xf_inv_point(X, x1);
        float t = path_entry(c, s, l0, l1, samples);
        if (t < best) {
            best = t;
            hit = s;
            l_hit = l0 + t * (l1 - l0);
        }
    }
    return hit;
}
__device__ f3 wall_friction(const Ctx& c, f3 v, f3 vb, f3 nn, float vn, int s) {
    if (c.friction == 0) return v;
    float mu = c.sh_mu[s];
    if (mu <= 0.0f) return v;
    f3 u = v - vb;
    f3 vt = u - dot(u, nn) * nn;
    float lt = length(vt);
    if (lt <= 0.0f) return v;
    return v - fmn(1.0f, mu * (-vn) / lt) * vt;
}

__device__ __forceinline__ void atomic_max_nonneg(float* p, float v) {
    atomicMax(reinterpret_cast<int*>(p), __float_as_int(v));
}

// block-wide sum / max of per-thread values (blockDim 256), result in every thread
__device__ __forceinline__ int block_sum_i32(int v) {
    __shared__ int red[4];
    for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int r = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return r;
}
__device__ __forceinline__ float block_max_f32(float v) {
    __shared__ float red[4];
    for (int s = 32; s >= 1; s >>= 1) v = fmx(v, __shfl_xor(v, s, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = fmx(fmx(red[0], red[1]), fmx(red[2], red[3]));
    __syncthreads();
    return r;
}

constexpr int LDS_SHAPES = 32;  // a world's shapes accumulated in LDS before one global atomic per shape and block

// speed_kernel (integrate.py:312-326) for one fluid particle: its speed, or 1e30 and a non-finite count
__device__ __forceinline__ void speed_stat(f3 v, f3 x, float& spmax, int& n_nonfinite) {
    float sp = length(v);
    if (isfinite(sp) && isfinite(x.x + x.y + x.z)) {
        spmax = fmx(spmax, sp);
    } else {
        n_nonfinite += 1;
        spmax = 1.0e30f;
    }
}

// shape meters (bodies.py:37-60, static shapes), advect_kernel, particle-wall advance, project_kernel and
// speed_kernel (integrate.py:181-326). Each particle reads only its own state and slots; per-world counters and the
// max speed are reduced in the block (a block is one world's chunk) and the shape meters in LDS, so global atomics
// are one per block instead of one per particle (int64 sums and maxima: the result is order independent).
__global__ void __launch_bounds__(256) k_advect_project(Ctx c, float dt) {
    CHUNK_PROLOGUE
    int w = c.ch_world[b_];
    __shared__ unsigned long long meter_lds[3 * LDS_SHAPES];
    int s0 = c.sh_start[w], ns = c.sh_start[w + 1] - s0;
    bool lds_meters = ns <= LDS_SHAPES;
    for (int t = threadIdx.x; t < 3 * LDS_SHAPES; t += blockDim.x) meter_lds[t] = 0ull;
    __syncthreads();
    int n_proj = 0, n_swept = 0, n_nonfinite = 0;
    float spmax = 0.0f;
    if (valid) {
        int64_t b4 = 4 * int64_t(i);
        if (!is_fluid(c, i)) {
            if (c.NW > 0 && c.wvol[i] > 0.0f) st3(c.q, i, Q(c, i) + dt * V(c, i));
        } else {
            for (int k = 0; k < 4 && !c.one_way; ++k) {  // bodies.accumulate_kernel meters (two-way only)
                int s = c.s_shape[b4 + k];
                if (s < 0 || s >= c.S) continue;
                f3 J = ld3(c.s_imp, b4 + k);
                if (J.x == 0.0f && J.y == 0.0f && J.z == 0.0f) continue;
                const double sc = 1125899906842624.0;
                unsigned long long v[3] = {(unsigned long long)(long long)llround(double(J.x) * sc),
                                           (unsigned long long)(long long)llround(double(J.y) * sc),
                                           (unsigned long long)(long long)llround(double(J.z) * sc)};
                int e = -1;
                if (lds_meters)
                    for (int t = 0; t < ns; ++t)
                        if (c.sh_ids[s0 + t] == s) {
                            e = t;
                            break;
                        }
                for (int d = 0; d < 3; ++d) {
                    if (e >= 0) atomicAdd(meter_lds + 3 * e + d, v[d]);
                    else atomicAdd(c.meter + 3 * s + d, v[d]);
                }
            }
            f3 v = V(c, i);
            f3 x = Q(c, i) + dt * v;
            int moved = 0;
            if (length(dt * v) > c.sweep_len) {
                f3 l;
                int s = sweep_entry(c, w, x - dt * v, x, l);
                if (s >= 0) {
                    Xf X = load_xf(c.sh_x + 7 * s);
                    float phi;
                    f3 gl;
                    shape_sdf(c.sh_geo[s], ld3(c.sh_scale, s), l, phi, gl);
                    f3 nn = normalize(xf_vector(X, gl));
                    x = xf_point(X, l) + fmx(c.eps - phi, 0.0f) * nn;
                    f3 vb{0.0f, 0.0f, 0.0f};
                    float vn = dot(v - vb, nn);
                    if (vn < 0.0f) {
                        v = v - vn * nn;
                        v = wall_friction(c, v, vb, nn, vn, s);
                    }
                    moved = 1;
                    n_swept = 1;
                }
            }
            int busy = 1, npass = 0;
            while (busy != 0 && npass < 4) {
                busy = 0;
                npass += 1;
                for (int b = 0; b < 4; ++b) {
                    int s = c.s_shape[b4 + b];
                    if (s == c.S && c.NW > 0) {  // pw_project (particle_walls.py:122-128)
                        f3 vb = ld3(c.s_vb, b4 + b), nb = ld3(c.s_nbl, b4 + b), xb = ld3(c.s_xbl, b4 + b);
                        float dd = dot(x - (xb + dt * vb), nb);
                        if (dd < c.eps) {
                            x = x + (c.eps - dd) * nb;
                            float vn = dot(v - vb, nb);
                            if (vn < 0.0f) v = v - vn * nb;
                            moved = 1;
                            busy = 1;
                        }
                    } else if (s >= 0 && s < c.S) {
                        Xf X = load_xf(c.sh_x + 7 * s);
                        f3 xl = xf_inv_point(X, x);
                        float phi;
                        f3 gl;
                        shape_sdf(c.sh_geo[s], ld3(c.sh_scale, s), xl, phi, gl);
                        if (phi < c.eps) {
                            f3 nn = normalize(xf_vector(X, gl));
                            float push = c.eps - phi;
                            if (c.s_phi[b4 + b] >= 0.0f) {
                                f3 ns_ = xf_vector(X, ld3(c.s_nbl, b4 + b));
                                float push_s = c.eps - dot(ns_, x - xf_point(X, ld3(c.s_xbl, b4 + b)));
                                if (push_s > 0.0f) {
                                    nn = ns_;
                                    push = push_s;
                                }
                            }
                            if (push > 0.0f) {
                                x = x + push * nn;
                                f3 vb{0.0f, 0.0f, 0.0f};
                                float vn = dot(v - vb, nn);
                                if (vn < 0.0f) {
                                    v = v - vn * nn;
                                    v = wall_friction(c, v, vb, nn, vn, s);
                                }
                                moved = 1;
                                busy = 1;
                            }
                        }
                    }
                }
            }
            st3(c.q, i, x);
            if (moved) {
                st3(c.qd, i, v);
                n_proj = 1;
            }
            if (!c.speed_later) speed_stat(v, x, spmax, n_nonfinite);
        }
    }
    __syncthreads();
    if (lds_meters)
        for (int t = threadIdx.x; t < 3 * ns; t += blockDim.x) {
            unsigned long long v = meter_lds[t];
            if (v != 0ull) atomicAdd(c.meter + 3 * c.sh_ids[s0 + t / 3] + t % 3, v);
        }
    n_proj = block_sum_i32(n_proj);
    n_swept = block_sum_i32(n_swept);
    n_nonfinite = block_sum_i32(n_nonfinite);
    spmax = block_max_f32(spmax);
    if (threadIdx.x == 0) {
        if (n_proj) atomicAdd(c.st_projected + w, n_proj);
        if (n_swept) atomicAdd(c.st_swept + w, n_swept);
        if (n_nonfinite) atomicAdd(c.st_nonfinite + w, n_nonfinite);
        atomic_max_nonneg(c.st_max_speed + w, spmax);
    }
}

// ----------------------------------------------------------------------------------------------- implicit viscosity
// Weiler et al. 2018 as in viscosity.py:1-220: A v = v*, (A v)_i = v_i - dt / rho_i sum_j coeff_ij (x_ij . (v_i - v_j))
// grad W_ij, a per-world CG with a 3x3 block-Jacobi preconditioner and fixed-order float64 dot products.
struct M33 {
    float m[9];  // row major
};
__device__ __forceinline__ f3 mul33(const M33& a, f3 v) {
    return f3{a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
              a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
// the inverse by cofactors over the determinant (expanded along the first row); a singular block gives zeros
__device__ __forceinline__ M33 inverse33(const M33& a) {
    const float* m = a.m;
    float c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    float det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    M33 r;
    if (det == 0.0f) {
        for (int k = 0; k < 9; ++k) r.m[k] = 0.0f;
        return r;
    }
    float inv = 1.0f / det;
    r.m[0] = c00 * inv;
    r.m[1] = (m[2] * m[7] - m[1] * m[8]) * inv;
    r.m[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
    r.m[3] = c01 * inv;
    r.m[4] = (m[0] * m[8] - m[2] * m[6]) * inv;
    r.m[5] = (m[2] * m[3] - m[0] * m[5]) * inv;
    r.m[6] = c02 * inv;
    r.m[7] = (m[1] * m[6] - m[0] * m[7]) * inv;
    r.m[8] = (m[0] * m[4] - m[1] * m[3]) * inv;
    return r;
}
// viscosity.py:39-47 _visc_coeff
__device__ __forceinline__ float visc_coeff(const Ctx& c, int i, int j, f3 xij) {
    float rho_j = c.ratio[j] * c.rho0[j];
    if (rho_j <= 0.0f) return 0.0f;
    float mu_ij = 0.5f * (c.mu[i] + c.mu[j]);
    return c.visc_d * mu_ij * (c.mass[j] / rho_j) / (dot(xij, xij) + 0.01f * c.h * c.h);
}

// wall_viscosity.py:96-105 _tangents: t1, t2 orthonormal to the unit vector n
__device__ __forceinline__ void wall_tangents(f3 n, f3& t1, f3& t2) {
    f3 a{1.0f, 0.0f, 0.0f};
    if (fabsf(n.x) > 0.6f) a = f3{0.0f, 1.0f, 0.0f};
    t1 = normalize(cross(n, a));
    t2 = cross(n, t1);
}

// wall_viscosity.py:108-150 slot_pair: the wall block M and w = M v_b of one slot (false for an empty slot or one
// whose term vanishes). The normal block comes from four samples around x_k = x_i - d n; the tangential block is the
// calibrated no-slip stress, faded out from 2r to 3r; v_b is the slot's wall velocity plus the particle's shift.
__device__ __forceinline__ bool wall_slot_pair(const Ctx& c, int64_t slot, f3 xi, f3 shift, M33& m, f3& w) {
    if (c.s_shape[slot] < 0) return false;
    const KC K = make_kc(c.h, c.kind);
    const float r = c.r, phi = c.s_phi[slot];
    const float d = fmx(phi + 0.5f * r, 2.0f * r);  // TimeStep.cpp:293 of SPlisHSPlasH
    const float wd = kernel_w(d, K);
    f3 n = xi - ld3(c.s_xb, slot);
    const float ln = length(n);
    float psi = c.s_psi[slot];
    if (phi > c.taper0) {  // the untapered half-space volume
        float v0, v1;
        halfspace_volume(phi + c.offset, c.h, 0.0f, c.kind, v0, v1);
        psi = v0 * c.amplitude;
    }
    if (wd <= 0.0f || psi <= 0.0f || ln < 1.0e-9f) return false;
    n = n / ln;
    const f3 xk = xi - d * n;
    f3 t1, t2;
    wall_tangents(n, t1, t2);
    const float vol = 0.25f * psi / wd;
#pragma unroll
    for (int k = 0; k < 9; ++k) m.m[k] = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float sg = 1.0f - 2.0f * float(q % 2);
        const f3 xs = xk + (sg * 0.5f * c.h) * (q < 2 ? t1 : t2);
        const f3 xis = xi - xs;
        const f3 g = kernel_grad(xis, K);
        const float cc = c.wall_coeff * vol / (dot(xis, xis) + 0.01f * c.h * c.h);
        const float gv[3] = {g.x, g.y, g.z}, xv[3] = {xis.x, xis.y, xis.z};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) m.m[3 * a + b] = m.m[3 * a + b] + cc * (gv[a] * xv[b]);
    }
    const float u = fcl((phi - 2.0f * r) / r, 0.0f, 1.0f);
    const float m_tt = -c.mu_b / (fmx(phi, r) * 2.0f * r) * (1.0f - u * u * (3.0f - 2.0f * u));
    const float mn = dot(n, mul33(m, n));
    const float nv[3] = {n.x, n.y, n.z};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float nn = nv[a] * nv[b];
            m.m[3 * a + b] = mn * nn + m_tt * ((a == b ? 1.0f : 0.0f) - nn);
        }
    w = mul33(m, ld3(c.s_vb, slot) + shift);
    return true;
}

// viscosity.py:235-243 pressure_guess_kernel (warm 1): the density solve's stored pressure as this step's guess
__global__ void __launch_bounds__(256) k_visc_guess(Ctx c, float dt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    c.vgk[i] = is_fluid(c, i) ? fmn(c.kd[i], 2.5e-4f) / (dt * dt) : 0.0f;
}

// viscosity.py:246-258 guess_shift_kernel: raw = -dt a, a the guess's pressure acceleration (pressure.py:264-277)
template <bool WIDE>
__global__ void __launch_bounds__(256) k_visc_raw(Ctx c, float dt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    if (!is_fluid(c, i)) {
        st3(c.vraw, i, f3{0.0f, 0.0f, 0.0f});
        return;
    }
    const KC K = make_kc(c.h, c.kind);
    const f3 xi = xyz(c.q4[i]);
    const float ki = c.vgk[i] / c.rho0[i];
    f3 a{0.0f, 0.0f, 0.0f};
    int cnt = c.count[i], base = c.ws[c.world[i]];
    for (int k = 0; k < cnt; ++k) {
        int j = nbr_at<WIDE>(c, base, i, k);
        f3 g = kernel_grad(xi - xyz(c.q4[j]), K);
        a = a - (c.mass[j] * (ki + c.vgk[j] / c.rho0[j])) * g;
    }
    const int64_t b4 = 4 * int64_t(i);
    for (int b = 0; b < 4; ++b)
        if (c.s_shape[b4 + b] >= 0) a = a - c.vgk[i] * ld3(c.s_gpsi, b4 + b);
    st3(c.vraw, i, (-dt) * a);
}

// viscosity.py:295-323 wall_shift_kernel: the Shepard average of raw over the liquid around a particle with a slot
template <bool WIDE>
__global__ void __launch_bounds__(256) k_visc_shift(Ctx c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const int64_t b4 = 4 * int64_t(i);
    bool wall = false;
    for (int b = 0; b < 4; ++b) wall = wall || c.s_shape[b4 + b] >= 0;
    if (!is_fluid(c, i) || !wall) {
        st3(c.vshift, i, f3{0.0f, 0.0f, 0.0f});
        return;
    }
    const KC K = make_kc(c.h, c.kind);
    const f3 xi = xyz(c.q4[i]);
    float tot = c.mass[i] / c.rho0[i] * kernel_w(0.0f, K);
    f3 acc = tot * ld3(c.vraw, i);
    int cnt = c.count[i], base = c.ws[c.world[i]];
    for (int k = 0; k < cnt; ++k) {
        int j = nbr_at<WIDE>(c, base, i, k);
        if (!is_fluid(c, j)) continue;
        float wj = c.mass[j] / c.rho0[j] * kernel_w(dist(xi - xyz(c.q4[j]), K), K);
        acc = acc + wj * ld3(c.vraw, j);
        tot = tot + wj;
    }
    st3(c.vshift, i, acc / tot);
}

// wall_viscosity.py:153-167 wall_terms_kernel: M_i = sum_k M_ik, w_i = sum_k w_ik
__global__ void __launch_bounds__(256) k_visc_wall_terms(Ctx c, int has_shift) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    M33 m;
#pragma unroll
    for (int k = 0; k < 9; ++k) m.m[k] = 0.0f;
    f3 w{0.0f, 0.0f, 0.0f};
    if (is_fluid(c, i)) {
        const f3 xi = xyz(c.q4[i]);
        const f3 sh = has_shift ? ld3(c.vshift, i) : f3{0.0f, 0.0f, 0.0f};
        for (int b = 0; b < 4; ++b) {
            M33 mk;
            f3 wk;
            if (!wall_slot_pair(c, 4 * int64_t(i) + b, xi, sh, mk, wk)) continue;
#pragma unroll
            for (int k = 0; k < 9; ++k) m.m[k] = m.m[k] + mk.m[k];
            w = w + wk;
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) c.wall_m[9 * int64_t(i) + k] = m.m[k];
    st3(c.wall_w, i, w);
}

// wall_viscosity.py:170-184 wall_reaction_kernel: each slot's wall takes -(m_i dt / rho_i)(M_ik v_i - w_ik)
__global__ void __launch_bounds__(256) k_visc_wall_reaction(Ctx c, float dt, int has_shift) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n || !is_fluid(c, i)) return;
    const f3 xi = xyz(c.q4[i]), vi = ld3(c.qd, i);
    const float rho_i = fmx(c.ratio[i] * c.rho0[i], 1.0e-6f);
    const float f = c.mass[i] * dt / rho_i;
    const f3 sh = has_shift ? ld3(c.vshift, i) : f3{0.0f, 0.0f, 0.0f};
    for (int b = 0; b < 4; ++b) {
        const int64_t slot = 4 * int64_t(i) + b;
        M33 mk;
        f3 wk;
        if (!wall_slot_pair(c, slot, xi, sh, mk, wk)) continue;  // an empty slot, or no term: no impulse
        st3(c.s_imp, slot, ld3(c.s_imp, slot) - f * (mul33(mk, vi) - wk));
    }
}

// viscosity.py:50-71 visc_apply_kernel: out = A vec and the chunk partial of vec . out
template <bool WIDE>
__global__ void __launch_bounds__(256) k_visc_apply(Ctx c, float dt, const float* vec, float* out) {
    CHUNK_PROLOGUE
    int w = c.ch_world[b_];
    if (c.v_conv[w] != 0) return;
    double px = 0.0;
    if (valid) {
        if (!is_fluid(c, i)) {
            st3(out, i, f3{0.0f, 0.0f, 0.0f});
        } else {
            const KC K = make_kc(c.h, c.kind);
            f3 xi = xyz(c.q4[i]), vi = ld3(vec, i), a{0.0f, 0.0f, 0.0f};
            int cnt = c.count[i], base = c.ws[c.world[i]];
            for (int k = 0; k < cnt; ++k) {
                int j = nbr_at<WIDE>(c, base, i, k);
                f3 xij = xi - xyz(c.q4[j]);
                a = a + (visc_coeff(c, i, j, xij) * dot(xij, vi - ld3(vec, j))) * kernel_grad(xij, K);
            }
            float rho_i = fmx(c.ratio[i] * c.rho0[i], 1.0e-6f);
            f3 r = vi - (dt / rho_i) * a;
            if (c.wall_coeff > 0.0f) {  // viscosity.py:69-70
                M33 wm;
#pragma unroll
                for (int k = 0; k < 9; ++k) wm.m[k] = c.wall_m[9 * int64_t(i) + k];
                r = r - (dt / rho_i) * mul33(wm, vi);
            }
            st3(out, i, r);
            px = double(dot(vi, r));
        }
    }
    chunk_reduce_x(c, b_, px, valid);
}

// viscosity.py:74-111 visc_init_kernel: the block preconditioner, r = b - A x, z = M^-1 r, p = z; partial (r.z, b.b)
template <bool WIDE>
__global__ void __launch_bounds__(256) k_visc_init(Ctx c, float dt) {
    CHUNK_PROLOGUE
    double px = 0.0, py = 0.0;
    if (valid) {
        if (!is_fluid(c, i)) {
            st3(c.vr, i, f3{0.0f, 0.0f, 0.0f});
            st3(c.vz, i, f3{0.0f, 0.0f, 0.0f});
            st3(c.vp, i, f3{0.0f, 0.0f, 0.0f});
        } else {
            const KC K = make_kc(c.h, c.kind);
            f3 xi = xyz(c.q4[i]);
            float rho_i = fmx(c.ratio[i] * c.rho0[i], 1.0e-6f);
            M33 blk{{1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f}};
            int cnt = c.count[i], base = c.ws[c.world[i]];
            for (int k = 0; k < cnt; ++k) {
                int j = nbr_at<WIDE>(c, base, i, k);
                f3 xij = xi - xyz(c.q4[j]);
                float s = (dt / rho_i) * visc_coeff(c, i, j, xij);
                f3 g = kernel_grad(xij, K);
                const float gv[3] = {g.x, g.y, g.z}, xv[3] = {xij.x, xij.y, xij.z};
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int q = 0; q < 3; ++q) blk.m[3 * r + q] = blk.m[3 * r + q] - s * (gv[r] * xv[q]);
            }
            f3 bi = ld3(c.vb, i);
            if (c.wall_coeff > 0.0f) {  // viscosity.py:97-99: the wall block and the moving walls' drag
                const float s = dt / rho_i;
#pragma unroll
                for (int k = 0; k < 9; ++k) blk.m[k] = blk.m[k] - s * c.wall_m[9 * int64_t(i) + k];
                bi = bi - s * ld3(c.wall_w, i);
                st3(c.vb, i, bi);
            }
            M33 mi = inverse33(blk);
#pragma unroll
            for (int k = 0; k < 9; ++k) c.vminv[9 * int64_t(i) + k] = mi.m[k];
            f3 ri = bi - ld3(c.vap, i);
            f3 zi = mul33(mi, ri);
            st3(c.vr, i, ri);
            st3(c.vz, i, zi);
            st3(c.vp, i, zi);
            px = double(dot(ri, zi));
            py = double(dot(bi, bi));
        }
    }
    chunk_reduce(c, b_, px, py, valid ? 0.0 : -3.4e38);
}

// viscosity.py:148-163 visc_update_kernel: x += a p, r -= a Ap, z = M^-1 r; partial (r.z, r.r)
__global__ void __launch_bounds__(256) k_visc_update(Ctx c) {
    CHUNK_PROLOGUE
    int w = c.ch_world[b_];
    if (c.v_conv[w] != 0) return;
    double px = 0.0, py = 0.0;
    if (valid && is_fluid(c, i)) {
        float a = float(c.v_step[w]);
        st3(c.vx, i, ld3(c.vx, i) + a * ld3(c.vp, i));
        f3 ri = ld3(c.vr, i) - a * ld3(c.vap, i);
        M33 mi;
#pragma unroll
        for (int k = 0; k < 9; ++k) mi.m[k] = c.vminv[9 * int64_t(i) + k];
        f3 zi = mul33(mi, ri);
        st3(c.vr, i, ri);
        st3(c.vz, i, zi);
        px = double(dot(ri, zi));
        py = double(dot(ri, ri));
    }
    chunk_reduce(c, b_, px, py, valid ? 0.0 : -3.4e38);
}

// viscosity.py:188-197 visc_direction_kernel; block 0 stores the unconverged-world count for the host
__global__ void __launch_bounds__(256) k_visc_direction(Ctx c, int* flag_out) {
    CHUNK_PROLOGUE
    if (flag_out && blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(flag_out, *c.n_unconv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    int w = c.ch_world[b_];
    if (c.v_conv[w] != 0 || !valid || !is_fluid(c, i)) return;
    st3(c.vp, i, ld3(c.vz, i) + float(c.v_step[w]) * ld3(c.vp, i));
}

// viscosity.py:200-204 visc_store_kernel
__global__ void __launch_bounds__(256) k_visc_store(Ctx c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < c.n && is_fluid(c, i)) st3(c.qd, i, ld3(c.vx, i));
}

// ----------------------------------------------------------------------------------------------- surface tension
// Continuum surface force with the Breinlinger contact angle (csf.py:63-249) and Akinci's pair force for small pieces
// (csf_sparse.py:51-88). Every pass reads positions, its inputs and the neighbour rows; the two velocity updates read
// no velocity of another particle, so they update in place.
constexpr float CSF_SURFACE_GRAD = 0.3f;  // |g| h above this marks a surface particle (csf.py:53)

__device__ __forceinline__ float csf_weight(float count, float lo, float hi) {  // csf_sparse.py:44-48
    float x = fcl((count - lo) / (hi - lo), 0.0f, 1.0f);
    return x * x * (3.0f - 2.0f * x);
}

// csf_sparse.py:51-62 max_count_kernel and csf.py:63-80 grad_kernel: the largest neighbour count around i, and the
// colour gradient with the wall slots counted as liquid
template <bool WIDE>
__global__ void __launch_bounds__(256) k_csf_grad(Ctx c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const KC K = make_kc(c.h, c.kind);
    const int cnt = c.count[i];
    const int base = c.world[i] >= 0 ? c.ws[c.world[i]] : 0;
    int s = cnt;
    for (int k = 0; k < cnt; ++k) s = max(s, c.count[nbr_at<WIDE>(c, base, i, k)]);
    c.t_mcount[i] = float(s);
    f3 g{0.0f, 0.0f, 0.0f};
    if (is_fluid(c, i)) {
        const f3 xi = xyz(c.q4[i]);
        const float inv = 1.0f / c.rho0[i];
        for (int k = 0; k < cnt; ++k) {
            int j = nbr_at<WIDE>(c, base, i, k);
            g = g + (c.mass[j] * inv) * kernel_grad(xi - xyz(c.q4[j]), K);
        }
        const int64_t b4 = 4 * int64_t(i);
        for (int b = 0; b < 4; ++b)
            if (c.s_shape[b4 + b] >= 0) g = g + 0.5f * ld3(c.s_gpsi, b4 + b);
    }
    st3(c.t_grad, i, g);
}

// csf.py:83-101 shepard_vec_kernel: the Shepard mean over the neighbours, self included
template <bool WIDE>
__global__ void __launch_bounds__(256) k_csf_shepard_vec(Ctx c, const float* vin, float* vout) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const f3 vi = ld3(vin, i);
    if (!is_fluid(c, i)) {
        st3(vout, i, vi);
        return;
    }
    const KC K = make_kc(c.h, c.kind);
    const f3 xi = xyz(c.q4[i]);
    const float w0 = kernel_w(0.0f, K);
    f3 s = w0 * vi;
    float den = w0;
    const int cnt = c.count[i], base = c.ws[c.world[i]];
    for (int k = 0; k < cnt; ++k) {
        int j = nbr_at<WIDE>(c, base, i, k);
        float w = kernel_w(dist(xi - xyz(c.q4[j]), K), K);
        s = s + w * ld3(vin, j);
        den = den + w;
    }
    st3(vout, i, s / den);
}

// csf.py:104-141 normal_kernel: the surface flag, the normal and the Breinlinger turn at a wall with a contact angle
__global__ void __launch_bounds__(256) k_csf_normal(Ctx c, const float* gdir) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const f3 z{0.0f, 0.0f, 0.0f};
    st3(c.t_nhat, i, z);
    st3(c.t_nwall, i, z);
    c.t_wslot[i] = -1;
    c.t_surf[i] = 0;
    const f3 gd = ld3(gdir, i);
    const float lg = length(gd);
    if (!is_fluid(c, i) || length(ld3(c.t_grad, i)) * c.h <= CSF_SURFACE_GRAD || lg <= 0.0f) return;
    f3 n = (-gd) / lg;
    const int64_t b4 = 4 * int64_t(i);
    int best = -1;
    float dmin = 1.0e30f;
    for (int b = 0; b < 4; ++b)
        if (c.s_shape[b4 + b] >= 0 && c.s_phi[b4 + b] < dmin) {
            dmin = c.s_phi[b4 + b];
            best = b;
        }
    if (best >= 0) {
        const int s = c.s_shape[b4 + best];
        if (s < c.S && c.sh_sin[s] >= 0.0f) {
            f3 nw = xyz(c.q4[i]) - ld3(c.s_xb, b4 + best);
            const float lw = length(nw);
            if (lw > 1.0e-9f) {
                nw = nw / lw;
                if (dmin < 0.0f) nw = -nw;
                const f3 t = n - dot(n, nw) * nw;
                const float lt = length(t);
                if (lt > 1.0e-6f) {
                    n = c.sh_cos[s] * nw + (c.sh_sin[s] / lt) * t;
                    st3(c.t_nwall, i, nw);
                    c.t_wslot[i] = best;
                }
            }
        }
    }
    st3(c.t_nhat, i, n);
    c.t_surf[i] = 1;
}

// csf.py:144-172 curvature_kernel: kappa = trace(B A^-1), the least-squares tangential gradient of n over the
// surface neighbours; A = sum w t t^T (+ 1e-2 trace(A) n n^T), B = sum w (P dn) t^T, t = P (x_j - x_i)
template <bool WIDE>
__global__ void __launch_bounds__(256) k_csf_curvature(Ctx c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    c.t_kraw[i] = 0.0f;
    if (c.t_surf[i] == 0) return;
    const KC K = make_kc(c.h, c.kind);
    const f3 xi = xyz(c.q4[i]), ni = ld3(c.t_nhat, i);
    const float nv[3] = {ni.x, ni.y, ni.z};
    M33 P, A, B;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            P.m[3 * a + b] = (a == b ? 1.0f : 0.0f) - nv[a] * nv[b];
            A.m[3 * a + b] = 0.0f;
            B.m[3 * a + b] = 0.0f;
        }
    int cnt = 0;
    const int nc = c.count[i], base = c.ws[c.world[i]];
    for (int k = 0; k < nc; ++k) {
        int j = nbr_at<WIDE>(c, base, i, k);
        if (c.t_surf[j] == 0) continue;
        const f3 xj = xyz(c.q4[j]);
        const float w = kernel_w(dist(xi - xj, K), K);
        const f3 t = mul33(P, xj - xi);
        const f3 d = mul33(P, ld3(c.t_nhat, j) - ni);
        const float tv[3] = {t.x, t.y, t.z}, dv[3] = {d.x, d.y, d.z};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                A.m[3 * a + b] = A.m[3 * a + b] + w * (tv[a] * tv[b]);
                B.m[3 * a + b] = B.m[3 * a + b] + w * (dv[a] * tv[b]);
            }
        cnt += 1;
    }
    const float scale = A.m[0] + A.m[4] + A.m[8];
    if (cnt < 4 || scale <= 0.0f) return;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) A.m[3 * a + b] = A.m[3 * a + b] + (1.0e-2f * scale) * (nv[a] * nv[b]);
    const M33 Ai = inverse33(A);
    float tr = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) tr = tr + (B.m[3 * a] * Ai.m[a] + B.m[3 * a + 1] * Ai.m[3 + a] + B.m[3 * a + 2] * Ai.m[6 + a]);
    c.t_kraw[i] = tr;
}

// csf.py:175-194 shepard_surface_kernel: the Shepard mean of the curvature over the surface neighbours
template <bool WIDE>
__global__ void __launch_bounds__(256) k_csf_shepard_surf(Ctx c, const float* kin, float* kout) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    kout[i] = kin[i];
    if (c.t_surf[i] == 0) return;
    const KC K = make_kc(c.h, c.kind);
    const f3 xi = xyz(c.q4[i]);
    const float w0 = kernel_w(0.0f, K);
    float num = w0 * kin[i], den = w0;
    const int cnt = c.count[i], base = c.ws[c.world[i]];
    for (int k = 0; k < cnt; ++k) {
        int j = nbr_at<WIDE>(c, base, i, k);
        if (c.t_surf[j] == 0) continue;
        float w = kernel_w(dist(xi - xyz(c.q4[j]), K), K);
        num = num + w * kin[j];
        den = den + w;
    }
    kout[i] = num / den;
}

// csf.py:197-226 csf_kernel: v_i += dt w_i a_i; at a contact line the part pointing away from the wall goes to the
// wall's slot impulse
__global__ void __launch_bounds__(256) k_csf_apply(Ctx c, const float* kin, float dt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    c.t_kappa[i] = 0.0f;
    const float wi = csf_weight(c.t_mcount[i], c.csf_lo, c.csf_hi);
    if (c.t_surf[i] == 0 || c.sigma[i] == 0.0f || wi <= 0.0f) return;
    float kap = kin[i];
    const float lg = length(ld3(c.t_grad, i));
    const float corr = 1.0f + 0.5f * kap * (c.ratio[i] - c.csf_cs) / lg;  // centre -> interface
    if (corr > 0.5f) kap = kap / corr;
    c.t_kappa[i] = kap;
    const float rho_i = fmx(c.ratio[i] * c.rho0[i], 1.0e-6f);
    f3 dv = (-(wi * dt * c.sigma[i] * kap * lg * c.csf_inv_i / rho_i)) * ld3(c.t_nhat, i);
    const int b = c.t_wslot[i];
    if (b >= 0) {
        const f3 nw = ld3(c.t_nwall, i);
        const float dn = dot(dv, nw);
        if (dn > 0.0f) {
            dv = dv - dn * nw;
            const int64_t slot = 4 * int64_t(i) + b;
            st3(c.s_imp, slot, ld3(c.s_imp, slot) + (c.mass[i] * dn) * nw);
        }
    }
    st3(c.qd, i, ld3(c.qd, i) + dv);
}

// tension.py:101-115 normals_kernel: Akinci's normal h sum_j (m_j / rho_j) grad W_ij
template <bool WIDE>
__global__ void __launch_bounds__(256) k_csf_akinci_normals(Ctx c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    f3 n{0.0f, 0.0f, 0.0f};
    if (is_fluid(c, i)) {
        const KC K = make_kc(c.h, c.kind);
        const f3 xi = xyz(c.q4[i]);
        const int cnt = c.count[i], base = c.ws[c.world[i]];
        for (int k = 0; k < cnt; ++k) {
            int j = nbr_at<WIDE>(c, base, i, k);
            const float rho_j = c.ratio[j] * c.rho0[j];
            if (rho_j > 0.0f) n = n + (c.mass[j] / rho_j) * kernel_grad(xi - xyz(c.q4[j]), K);
        }
        n = c.h * n;
    }
    st3(c.t_anorm, i, n);
}

// tension.py:81-98 cohesion_c and curvature_weight
__device__ __forceinline__ float cohesion_c(float r, float h, float k) {
    if (r > h || r <= 0.0f) return 0.0f;
    const float s = (h - r) * (h - r) * (h - r) * r * r * r;
    if (r > 0.5f * h) return k * s;
    const float h3 = h * h * h;
    return k * (2.0f * s - h3 * h3 / 64.0f);
}

// csf_sparse.py:65-88 sparse_pair_kernel: v_i += dt sum_j c_ij K_ij gamma_ij [curvature + cohesion] on small pieces
template <bool WIDE>
__global__ void __launch_bounds__(256) k_csf_pairs(Ctx c, float dt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n || !is_fluid(c, i) || c.gamma[i] == 0.0f) return;
    const f3 xi = xyz(c.q4[i]), ni = ld3(c.t_anorm, i);
    const float rho_i = c.ratio[i] * c.rho0[i];
    const float wi = csf_weight(c.t_mcount[i], c.csf_lo, c.csf_hi);
    f3 a{0.0f, 0.0f, 0.0f};
    const int cnt = c.count[i], base = c.ws[c.world[i]];
    for (int k = 0; k < cnt; ++k) {
        int j = nbr_at<WIDE>(c, base, i, k);
        const float cc = 1.0f - 0.5f * (wi + csf_weight(c.t_mcount[j], c.csf_lo, c.csf_hi));
        const float g = 0.5f * (c.gamma[i] + c.gamma[j]);
        const float rho_j = c.ratio[j] * c.rho0[j];
        if (cc <= 0.0f || g == 0.0f || rho_i + rho_j <= 0.0f) continue;
        const f3 rv = xi - xyz(c.q4[j]);
        const float r = length(rv);
        f3 acc = (4.0f * fmx(1.0f - r / c.h, 0.0f)) * (ni - ld3(c.t_anorm, j));
        if (r > 1.0e-6f * c.h) acc = acc + (c.mass[j] * cohesion_c(r, c.h, c.csf_kcoh) / r) * rv;
        a = a - (cc * 2.0f * c.rho0[i] / (rho_i + rho_j) * g) * acc;
    }
    st3(c.qd, i, ld3(c.qd, i) + dt * a);
}

// ----------------------------------------------------------------------------------------------- emitters
// Box emitters, sinks, the escape box and the int64 ledgers (emitters.py:32-255, runtime.py:197-409). The tables are
// device copies of the host's emitter table; per-world ledgers are int64 fixed point, so their atomic sums do not
// depend on the order of the additions.
struct EmitCtx {
    int E, has_bounds;
    f3 lo, hi;                    // escape box (config.world_bounds)
    const int* e_world;           // [E]
    const float* e_xform;         // [E*7] region transform
    const float* e_vel;           // [E*3]
    const float* e_interval;      // [E] seconds per layer
    const int* off_start;         // [E+1] CSR into off_local
    const float* off_local;       // [*3] layer lattice in the region frame
    const int* pool_start;        // [E+1] CSR into pool_ids
    const int* pool_ids;          // spare ids in emission order
    float* t_acc;                 // [E] layer clock
    int* cursor;                  // [E] spares used
    int* enabled;                 // [E]
    const int* sink_start;        // [W+1] CSR of the sinks of each world
    const float* s_xform;         // [*7]
    const float* s_half;          // [*3]
    long long *mass0, *mass_in, *mass_out, *mass_act, *flux_e, *flux_g;  // [W] fixed point
    int *escaped, *emit_overflow;                                         // [W]
};

// kernels.py:142-144 to_fixed at 2^40 kg; bodies.py:31-32 _fx
__device__ __forceinline__ long long to_fixed_scaled(double x, double scale) { return llround(x * scale); }
__device__ __forceinline__ long long to_fixed(float m) { return llround(double(m) * 1099511627776.0); }
__device__ __forceinline__ void add_i64(long long* p, long long v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

// emitters.py:82-89 flux_add: the energy m (v^2/2 - g.x) [2^50 J] and the slope m g_z [2^40 N] a particle brings in
// (sign +1) or takes out (-1)
__device__ __forceinline__ void flux_add(const EmitCtx& e, int w, float m, f3 v, f3 x, f3 g, float sign) {
    float en = sign * m * (0.5f * dot(v, v) - dot(g, x));
    add_i64(e.flux_e + w, to_fixed_scaled(double(en), 1125899906842624.0));
    add_i64(e.flux_g + w, to_fixed_scaled(double(sign * m * g.z), 1099511627776.0));
}

// emitters.py:127-130 _in_box
__device__ __forceinline__ bool in_box(const Xf& X, f3 half, f3 x) {
    f3 xl = xf_inv_point(X, x);
    return fabsf(xl.x) <= half.x && fabsf(xl.y) <= half.y && fabsf(xl.z) <= half.z;
}

// emitters.py:133-165 retire_kernel (sinks of the particle's world, then the escape box) and, for the particles that
// stay, speed_kernel (integrate.py:312-326), which the reference runs after emission; emitted particles add their
// own speed in k_emit. One block per chunk (one world): the speed statistics are reduced in the block.
__global__ void __launch_bounds__(256) k_retire(Ctx c, EmitCtx e) {
    CHUNK_PROLOGUE
    int w = c.ch_world[b_];
    int n_nonfinite = 0;
    float spmax = 0.0f;
    if (valid && is_fluid(c, i)) {
        f3 x = Q(c, i);
        int gone = 0;
        for (int s = e.sink_start[w]; s < e.sink_start[w + 1]; ++s)
            if (in_box(load_xf(e.s_xform + 7 * s), ld3(e.s_half, s), x)) gone = 1;
        if (e.has_bounds != 0 && gone == 0) {
            bool out = x.x < e.lo.x || x.y < e.lo.y || x.z < e.lo.z || x.x > e.hi.x || x.y > e.hi.y || x.z > e.hi.z;
            if (out || !isfinite(x.x + x.y + x.z)) {
                gone = 2;
                atomicAdd(e.escaped + w, 1);
            }
        }
        if (gone != 0) {
            f3 v = V(c, i);
            float m = c.mass[i];
            if (isfinite(x.x + x.y + x.z + v.x + v.y + v.z)) flux_add(e, w, m, v, x, ld3(c.gravity, w), -1.0f);
            c.flags[i] = c.flags[i] & ~ACTIVE;
            c.phase[i] = int8_t(-2);  // retired: never reused by an emitter
            st3(c.qd, i, f3{0.0f, 0.0f, 0.0f});
            add_i64(e.mass_out + w, to_fixed(m));
        } else {
            speed_stat(V(c, i), x, spmax, n_nonfinite);
        }
    }
    n_nonfinite = block_sum_i32(n_nonfinite);
    spmax = block_max_f32(spmax);
    if (threadIdx.x == 0) {
        if (n_nonfinite) atomicAdd(c.st_nonfinite + w, n_nonfinite);
        atomic_max_nonneg(c.st_max_speed + w, spmax);
    }
}

// emitters.py:67-79 emit_clock_kernel, 92-117 emit_kernel and 120-124 advance_cursor_kernel for one emitter per
// block: the clock, then a layer from the emitter's pool (whole spares in pool order; the ones past the pool count
// as overflow), then the cursor. Emitted particles enter speed_kernel's statistics here.
__global__ void __launch_bounds__(256) k_emit(Ctx c, EmitCtx e, float dt) {
    const int em = blockIdx.x;
    __shared__ int now;
    if (threadIdx.x == 0) {
        int go = 0;
        if (e.enabled[em] != 0) {
            float t = e.t_acc[em] + dt;
            if (t >= e.e_interval[em]) {
                t -= e.e_interval[em];
                go = 1;
            }
            e.t_acc[em] = t;
        }
        now = go;
    }
    __syncthreads();
    if (now == 0) return;
    const int w = e.e_world[em];
    const int o0 = e.off_start[em], nl = e.off_start[em + 1] - o0;
    const int p0 = e.pool_start[em], psize = e.pool_start[em + 1] - p0;
    const int cur = e.cursor[em];
    const Xf X = load_xf(e.e_xform + 7 * em);
    const f3 vel = ld3(e.e_vel, em);
    const f3 g = ld3(c.gravity, w);
    int over = 0, n_nonfinite = 0;
    float spmax = 0.0f;
    for (int r = threadIdx.x; r < nl; r += blockDim.x) {
        int slot = cur + r;
        if (slot >= psize) {
            over += 1;
            continue;
        }
        int i = e.pool_ids[p0 + slot];
        f3 x = xf_point(X, ld3(e.off_local, o0 + r));
        st3(c.q, i, x);
        st3(c.qd, i, vel);
        c.flags[i] = c.flags[i] | ACTIVE;
        c.phase[i] = int8_t(1);
        float m = c.mass[i];
        add_i64(e.mass_in + w, to_fixed(m));
        flux_add(e, w, m, vel, x, g, 1.0f);
        if (c.speed_later) speed_stat(vel, x, spmax, n_nonfinite);
    }
    over = block_sum_i32(over);
    n_nonfinite = block_sum_i32(n_nonfinite);
    spmax = block_max_f32(spmax);  // every thread has passed its reads of the cursor
    if (threadIdx.x == 0) {
        if (over) atomicAdd(e.emit_overflow + w, over);
        if (n_nonfinite) atomicAdd(c.st_nonfinite + w, n_nonfinite);
        if (c.speed_later && spmax > 0.0f) atomic_max_nonneg(c.st_max_speed + w, spmax);
        e.cursor[em] = min(cur + nl, psize);
    }
}

// ports.py:66-93 _take_kernel (retire): the masked liquid particles leave for good, their mass and energy out of the
// ledgers as at a sink
__global__ void __launch_bounds__(256) k_retire_mask(Ctx c, EmitCtx e, const int* mask) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n || mask[i] == 0 || !is_fluid(c, i)) return;
    int w = c.world[i];
    f3 x = Q(c, i), v = V(c, i);
    float m = c.mass[i];
    if (isfinite(x.x + x.y + x.z + v.x + v.y + v.z)) flux_add(e, w, m, v, x, ld3(c.gravity, w), -1.0f);
    c.flags[i] = c.flags[i] & ~ACTIVE;
    st3(c.qd, i, f3{0.0f, 0.0f, 0.0f});
    add_i64(e.mass_out + w, to_fixed(m));
    c.phase[i] = int8_t(-2);
}

// emitters.py:168-172 active_mass_kernel: the fixed-point mass of each world's active liquid (block sums, then one
// atomic per block)
__global__ void __launch_bounds__(256) k_active_mass(Ctx c, long long* out) {
    CHUNK_PROLOGUE
    int w = c.ch_world[b_];
    __shared__ unsigned long long acc;
    if (threadIdx.x == 0) acc = 0ull;
    __syncthreads();
    if (valid && is_fluid(c, i)) atomicAdd(&acc, static_cast<unsigned long long>(to_fixed(c.mass[i])));
    __syncthreads();
    if (threadIdx.x == 0 && acc != 0ull) atomicAdd(reinterpret_cast<unsigned long long*>(out) + w, acc);
}

// ----------------------------------------------------------------------------------------------- learned warm start
// The network predicts the correction y to the rule in asinh(kappa / ks) units: kappa0 = ks sinh(rule + y), so a
// zero output reproduces the reference's warm start.
__global__ void __launch_bounds__(256) k_warm_from_mlp(Ctx c, float dt) {
    CHUNK_PROLOGUE
    if (!valid || (c.learn_gate && !c.learn_on[c.ch_world[b_]])) return;
    const float* F = c.feat + 16 * int64_t(i);
    float y0 = c.mlp_out[2 * int64_t(i)], y1 = c.mlp_out[2 * int64_t(i) + 1];
    float kv = c.kscale * sinhf(F[13] + y0), kd = c.kscale * sinhf(F[14] + y1);
    if (!isfinite(kv)) kv = 0.0f;
    if (!isfinite(kd)) kd = 0.0f;
    c.warm_k[i] = fmn(kv, 0.5f / dt);
    c.warm_k[int64_t(c.n) + i] = fmx(fmn(kd, 2.5e-4f / (dt * dt)), 0.0f);
}

// The learned start's per-world gate (DFSPH_OPT_LEARNED_GATE): a world takes the network's start at a step when the
// last step's solve that the network feeds (fed: bit 0 divergence, bit 1 density) needed more than min_iterations
// iterations. A solve held at the min_iterations floor cannot save an iteration, so there the network would only
// cost; such a world (and every world at its first step) starts from the reference's rule.
__global__ void k_learn_flags(Ctx c, int fed, int min_it) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W) return;
    c.learn_on[w] = ((fed & 1) && c.it_div[w] > min_it) || ((fed & 2) && c.it_den[w] > min_it) ? 1 : 0;
}
// per 64-row tile of the network's input rows: 1 when a particle of the tile belongs to a world with the learned start
__global__ void __launch_bounds__(256) k_learn_tiles(Ctx c, uint8_t* tile_on) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int w = i < c.n ? c.world[i] : -1;
    const bool on = w >= 0 && c.learn_on[w] != 0;
    const unsigned long long any = __ballot(on);
    if ((threadIdx.x & 63) == 0 && i < c.n) tile_on[i >> 6] = any != 0ull ? 1 : 0;
}

// training targets of the step just taken, [n][2]: asinh(converged kappa / ks) of both solves minus the rule's
// initial value logged with the step's features (the correction the network learns)
__global__ void k_targets(Ctx c, float dt, float* T) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const float* F = c.feat + 16 * int64_t(i);
    T[2 * int64_t(i)] = asinhf(c.kv[i] / dt / c.kscale) - F[13];
    T[2 * int64_t(i) + 1] = asinhf(c.kd[i] / (dt * dt) / c.kscale) - F[14];
}

// ----------------------------------------------------------------------------------------------- profiling
// Optional per-launch timing of the main kernels: the start / stop events ride in the kernel's own dispatch
// (hipExtLaunchKernelGGL), so a duration is the kernel's execution alone, as rocprofv3's kernel trace reports it,
// never the gap before a starved launch. Finished pairs are summed and recycled at the next step, which keeps the
// event pool bounded over any number of steps.
struct Prof {
    bool on = false;
    uint32_t mask = 0;  // kernel ids timed
    int sample = 1;     // time a pseudo-random 1 / sample of those launches (host cost of a timed launch)
    uint64_t launches[KID_COUNT] = {};
    struct Rec {
        int kid;
        hipEvent_t a, b;
    };
    std::deque<Rec> pending;
    std::vector<hipEvent_t> pool;
    double total_ms[KID_COUNT] = {};
    int count[KID_COUNT] = {};
    unsigned long long* d_bytes = nullptr;  // [2][n_chunks][4] algorithmic bytes counted by k_accel / k_apply
};

// ----------------------------------------------------------------------------------------------- handle
struct GraphKey {
    int kind, mode, every, pair_cache;
    uint32_t dt_bits;
    const void *q, *qd, *f, *ld, *ln, *mlp, *lf, *lt;
    bool operator<(const GraphKey& o) const {
        return std::tie(kind, mode, every, pair_cache, dt_bits, q, qd, f, ld, ln, mlp, lf, lt) <
               std::tie(o.kind, o.mode, o.every, o.pair_cache, o.dt_bits, o.q, o.qd, o.f, o.ld, o.ln, o.mlp, o.lf,
                        o.lt);
    }
};

struct Handle {
    DfsphParams P;
    Ctx c;
    int n_chunks = 0, n_scan_blocks = 0;
    int64_t nbr_bytes = 0, nbr_hi_bytes = 0;
    uint16_t* nbr_view = nullptr;  // the rows in the reference's layout [cap][n], made on request (rows_view)
    uint8_t* nbr_hi_view = nullptr;
    int* h_unconv = nullptr;  // pinned [2]: unconverged-world counts of the last two iteration batches
    int* d_unconv = nullptr;  // the same memory as the device addresses it
    hipEvent_t ev_unconv[2] = {nullptr, nullptr};
    std::vector<void*> owned;
    float* qd_saved = nullptr;
    float* learned_div = nullptr;
    float* learned_den = nullptr;
    int check_every = 1;
    int device_loop = 0;                // DFSPH_OPT_DEVICE_LOOP: the CG iterations as one cooperative launch (k_cg_loop)
    int loop_grid[2] = {0, 0};          // its grid (co-resident blocks, a multiple of 8) per row width, 0: not yet known
    Prof prof;
    bool use_graphs = false;  // measured: eager launches are faster than graph replay for this step (bench.py --graphs)
    const float* mlp_params = nullptr;  // learned warm start on when set
    int learned_solves = 3;             // DFSPH_OPT_LEARNED_SOLVES: bit 0 divergence, bit 1 density
    int learned_gate = 1;               // DFSPH_OPT_LEARNED_GATE
    uint8_t* learn_tiles = nullptr;     // [(n + 63) / 64] tiles of the network's rows with a gated-on world
    int build_block = 128;              // DFSPH_OPT_BUILD_BLOCK
    int build_staged = 0;               // DFSPH_OPT_BUILD_STAGED
    int packed_ok = 0;                  // every world's SPH particles share one mass: DFSPH_OPT_PACKED_ACCEL allowed
    float* log_feat = nullptr;          // when set, the step copies its features [n][16] here ...
    float* log_targets = nullptr;       // ... and its converged-kappa targets [n][2] here
    hipStream_t cap_stream = nullptr;
    std::map<GraphKey, hipGraphExec_t> graphs;
    EmitCtx em{};                      // emitter, sink and escape-box tables and the per-world ledgers
    int emit_active = 0;               // any emitter, sink or escape box: the step ends with the emitter pass
    int sink_count = 0;
    std::vector<void*> emit_owned;     // the current tables (replaced by dfsph_set_emitters)
    int n_wall = 0;                    // particles with a wall volume (the exchange window's rows)
};

// an event pair (and, for k_accel / k_apply, the byte counters) for a timed launch of kernel kid, or false when
// this launch is not timed; the 1 / sample choice hashes the launch count, so it does not alias with the fixed
// launch pattern of a solver iteration
bool prof_begin(Handle* H, int kid, hipEvent_t* a, hipEvent_t* b, unsigned long long** bytes) {
    Prof& P = H->prof;
    if (!P.on || !((P.mask >> kid) & 1u)) return false;
    uint64_t z = (P.launches[kid]++ + (uint64_t(kid) << 48)) * 0x9E3779B97F4A7C15ull;  // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    if (P.sample > 1 && z % uint64_t(P.sample) != 0) return false;
    hipEvent_t e[2];
    for (hipEvent_t& x : e) {
        if (!P.pool.empty()) {
            x = P.pool.back();
            P.pool.pop_back();
        } else if (hipEventCreateWithFlags(&x, hipEventDisableSystemFence) != hipSuccess) {
            // timing-only events: no system-scope fence (an L2 writeback at every timed launch would slow the
            // kernels being timed, and the ones after them)
            return false;
        }
    }
    *a = e[0];
    *b = e[1];
    *bytes = (kid == KID_ACCEL || kid == KID_APPLY) ? P.d_bytes : nullptr;
    P.pending.push_back(Prof::Rec{kid, e[0], e[1]});
    return true;
}
// sum the finished timed launches (all of them with `wait`), returning their events to the pool
int prof_harvest(Handle* H, bool wait) {
    Prof& P = H->prof;
    while (!P.pending.empty()) {
        Prof::Rec r = P.pending.front();
        if (wait) {
            HIPCHK(hipEventSynchronize(r.b));
        } else if (hipEventQuery(r.b) != hipSuccess) {
            break;
        }
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        P.total_ms[r.kid] += ms;
        P.count[r.kid] += 1;
        P.pool.push_back(r.a);
        P.pool.push_back(r.b);
        P.pending.pop_front();
    }
    return DFSPH_OK;
}
// launch `kernel` (first argument the Ctx `ctx`), timed as kernel id `kid` when profiling samples it: the timed
// launch carries its events in its dispatch and a Ctx copy whose prof_bytes points at the byte counters
#define LAUNCH_T(kid, kernel, grid, block, ctx, ...)                                                       \
    do {                                                                                                   \
        hipEvent_t ea_, eb_;                                                                               \
        unsigned long long* pb_ = nullptr;                                                                 \
        if (prof_begin(H, kid, &ea_, &eb_, &pb_)) {                                                        \
            Ctx ct_ = (ctx);                                                                               \
            ct_.prof_bytes = pb_;                                                                          \
            hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, ea_, eb_, 0u, ct_, ##__VA_ARGS__); \
        } else {                                                                                           \
            kernel<<<(grid), (block), 0, st>>>((ctx), ##__VA_ARGS__);                                      \
        }                                                                                                  \
    } while (0)
#define LAUNCH_WT(kid, kernel, grid, block, ctx, ...)                                    \
    do {                                                                                 \
        if ((ctx).wide) LAUNCH_T(kid, kernel<true>, grid, block, ctx, ##__VA_ARGS__);    \
        else LAUNCH_T(kid, kernel<false>, grid, block, ctx, ##__VA_ARGS__);              \
    } while (0)

// the pressure-acceleration and operator-apply passes, LDS-staged when DFSPH_OPT_HALO is on
#define LAUNCH_ACCEL(ctx, ...)                                                                   \
    do {                                                                                         \
        if (!(ctx).use_halo) LAUNCH_WT(KID_ACCEL, k_accel, H->n_chunks, 256, ctx, __VA_ARGS__);  \
        else if ((ctx).packed) LAUNCH_WT(KID_ACCEL, k_accel_hp, H->n_chunks, 256, ctx, __VA_ARGS__); \
        else LAUNCH_WT(KID_ACCEL, k_accel_hn, H->n_chunks, 256, ctx, __VA_ARGS__);               \
    } while (0)
#define LAUNCH_APPLY(ctx, ...)                                                                   \
    do {                                                                                         \
        if (!(ctx).use_halo) LAUNCH_WT(KID_APPLY, k_apply, H->n_chunks, 256, ctx, __VA_ARGS__);  \
        else if ((ctx).packed) LAUNCH_WT(KID_APPLY, k_apply_hp, H->n_chunks, 256, ctx, __VA_ARGS__); \
        else LAUNCH_WT(KID_APPLY, k_apply_hn, H->n_chunks, 256, ctx, __VA_ARGS__);               \
    } while (0)

template <class T>
int dalloc(Handle* H, T** p, int64_t count) {
    void* q = nullptr;
    size_t bytes = size_t(std::max<int64_t>(count, 1)) * sizeof(T);
    HIPCHK(hipMalloc(&q, bytes));
    HIPCHK(hipMemset(q, 0, bytes));
    H->owned.push_back(q);
    *p = static_cast<T*>(q);
    return DFSPH_OK;
}
template <class T>
int dupload(Handle* H, T** p, const T* host, int64_t count) {
    int rc = dalloc(H, p, count);
    if (rc) return rc;
    if (count > 0) HIPCHK(hipMemcpy(*p, host, size_t(count) * sizeof(T), hipMemcpyHostToDevice));
    return DFSPH_OK;
}

inline int blocks_for(int64_t n, int b = 256) { return int(std::max<int64_t>((n + b - 1) / b, 1)); }

#define LAUNCH_CHECK()                                                                   \
    do {                                                                                 \
        hipError_t e_ = hipGetLastError();                                               \
        if (e_ != hipSuccess) return fail(DFSPH_EDEVICE, std::string("launch: ") + hipGetErrorString(e_)); \
    } while (0)

Handle* as_handle(void* h) { return static_cast<Handle*>(h); }

// ----------------------------------------------------------------------------------------------- phases (host)
template <bool WIDE, int BB>
void launch_build_lists_bb(Handle* H, hipStream_t st, bool sup) {
    const Ctx& c = H->c;
    auto kf = sup ? k_build_lists<WIDE, BB, true> : k_build_lists<WIDE, BB, false>;
    hipEvent_t ea, eb;
    unsigned long long* pb = nullptr;
    int grid = blocks_for(c.n, BB);
    if (prof_begin(H, KID_BUILD, &ea, &eb, &pb)) hipExtLaunchKernelGGL(kf, dim3(grid), dim3(BB), 0, st, ea, eb, 0u, c);
    else hipLaunchKernelGGL(kf, dim3(grid), dim3(BB), 0, st, c);
}
template <bool WIDE, int BB>
void launch_build_staged_bb(Handle* H, hipStream_t st) {
    const Ctx& c = H->c;
    auto kf = k_build_lists_staged<WIDE, BB>;
    hipEvent_t ea, eb;
    unsigned long long* pb = nullptr;
    const int grid = std::max(H->n_chunks, 1) * (CHUNK / BB);
    if (prof_begin(H, KID_BUILD, &ea, &eb, &pb)) hipExtLaunchKernelGGL(kf, dim3(grid), dim3(BB), 0, st, ea, eb, 0u, c);
    else hipLaunchKernelGGL(kf, dim3(grid), dim3(BB), 0, st, c);
}
// the list build with the tuned block size (64, 128 or 256; 256 only with uint16 rows: 64 KB of LDS otherwise); the
// plain rows with the staged build (64 or 128 threads) when DFSPH_OPT_BUILD_STAGED is on
void launch_build_lists(Handle* H, hipStream_t st, bool sup = false) {
    if (!sup && H->build_staged) {
        if (H->c.wide) {
            if (H->build_block == 64) launch_build_staged_bb<true, 64>(H, st);
            else launch_build_staged_bb<true, 128>(H, st);
        } else {
            if (H->build_block == 64) launch_build_staged_bb<false, 64>(H, st);
            else launch_build_staged_bb<false, 128>(H, st);
        }
        return;
    }
    if (H->c.wide) {
        if (H->build_block == 64) launch_build_lists_bb<true, 64>(H, st, sup);
        else launch_build_lists_bb<true, 128>(H, st, sup);
    } else {
        if (H->build_block == 64) launch_build_lists_bb<false, 64>(H, st, sup);
        else if (H->build_block == 256) launch_build_lists_bb<false, 256>(H, st, sup);
        else launch_build_lists_bb<false, 128>(H, st, sup);
    }
}

int phase_neighbors(Handle* H, hipStream_t st) {
    Ctx& c = H->c;
    if (c.n == 0) return DFSPH_OK;
    k_fill_i32<<<std::min(blocks_for(int64_t(c.n_cells) + 1), 4096), 256, 0, st>>>(c.cell_count, 0,
                                                                                   int64_t(c.n_cells) + 1);
    k_fill_i32<<<blocks_for(c.n_chunks), 256, 0, st>>>(c.ch_live, 0, c.n_chunks);
    k_groups_hist<<<blocks_for(c.n), 256, 0, st>>>(c);
    k_scan_reduce<<<H->n_scan_blocks, 256, 0, st>>>(c.cell_count, c.n_cells, c.scan_part);
    k_scan_top<<<1, 256, 0, st>>>(c.scan_part, H->n_scan_blocks, c.cell_start, c.n_cells);
    k_scan_apply<<<H->n_scan_blocks, 256, 0, st>>>(c.cell_count, c.n_cells, c.scan_part, c.cell_start);
    k_scatter<<<blocks_for(c.n), 256, 0, st>>>(c);
    k_place_cells<<<blocks_for(c.n), 256, 0, st>>>(c);  // ascending ids in every cell
    if (c.NW > 0) {
        int64_t nw_cells = int64_t(c.n_cells) - c.wall_first;
        k_fill_u8<<<std::min(blocks_for(nw_cells), 4096), 256, 0, st>>>(c.wall_near, nw_cells);
        k_mark_wall_cells<<<blocks_for(c.n), 256, 0, st>>>(c);
    }
    if (c.skin > 0.0f) {  // neighbors_skin.py:179-203: flags, the flagged worlds' superset rows, then the filter
        k_skin_seed<<<blocks_for(c.W), 256, 0, st>>>(c);
        k_skin_check<<<blocks_for(c.n), 256, 0, st>>>(c);
        launch_build_lists(H, st, true);
        if (c.wide) k_skin_filter<true><<<blocks_for(c.n), 256, 0, st>>>(c);
        else k_skin_filter<false><<<blocks_for(c.n), 256, 0, st>>>(c);
    } else {
        launch_build_lists(H, st);
    }
    if (c.use_halo) {
        if (c.wide) k_halo_build<true><<<H->n_chunks, 256, 0, st>>>(c);
        else k_halo_build<false><<<H->n_chunks, 256, 0, st>>>(c);
    }
    LAUNCH_CHECK();
    return DFSPH_OK;
}

int phase_slots(Handle* H, hipStream_t st) {
    Ctx& c = H->c;
    if (c.n == 0) return DFSPH_OK;
    LAUNCH_T(KID_SLOTS, k_slots, blocks_for(c.n), 256, c);
    LAUNCH_CHECK();
    return DFSPH_OK;
}

// density and alpha; with `features` (the step's dt > 0) also the learned warm start's features
int phase_density(Handle* H, hipStream_t st, float features_dt = 0.0f) {
    Ctx& c = H->c;
    if (c.n == 0) return DFSPH_OK;
    if (features_dt > 0.0f) {
        LAUNCH_WT(KID_DENSITY, k_density_features, H->n_chunks, 256, c, features_dt, c.learn_gate);
        if (c.learn_gate) LAUNCH_WT(KID_DENSITY, k_density_alpha, H->n_chunks, 256, c, 1);
    } else {
        LAUNCH_WT(KID_DENSITY, k_density_alpha, H->n_chunks, 256, c, 0);
    }
    k_world<<<c.W, 256, 0, st>>>(c, W_NACTIVE, StopSpec{});
    LAUNCH_CHECK();
    return DFSPH_OK;
}

int phase_flow(Handle* H, hipStream_t st) {
    Ctx& c = H->c;
    if (!(H->P.rest_jitter_speed > 0.0) || !(H->P.rest_jitter_flow_ratio > 0.0) || c.n == 0) return DFSPH_OK;
    k_flow_terms<<<H->n_chunks, 256, 0, st>>>(c, 0);
    k_world<<<c.W, 256, 0, st>>>(c, W_FLOW_Z, StopSpec{});
    k_flow_terms<<<H->n_chunks, 256, 0, st>>>(c, 1);
    k_world<<<c.W, 256, 0, st>>>(c, W_FLOW_XY, StopSpec{});
    LAUNCH_CHECK();
    return DFSPH_OK;
}

// The device-side loop replacing wp.capture_while: batch b leaves the unconverged-world count in pinned slot b & 1
// (stored by its last kernel, or copied where the batch is a replayed graph or a Jacobi batch) and records that
// slot's event; the host enqueues batch b + 1 before it waits on batch b's event, so the GPU never idles on the
// host. The batch enqueued after the last unconverged world converged is a no-op (every kernel of a converged world
// exits at its first instruction).
int post_unconv(Handle* H, hipStream_t st, int slot, bool stored) {
    if (!stored)
        HIPCHK(hipMemcpyAsync(H->h_unconv + slot, H->c.n_unconv, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(H->ev_unconv[slot], st));
    return DFSPH_OK;
}
int wait_unconv(Handle* H, int slot, int* out) {
    HIPCHK(hipEventSynchronize(H->ev_unconv[slot]));
    *out = H->h_unconv[slot];
    return DFSPH_OK;
}

struct SolveOverride {
    double tol = -1.0, err_factor = -1.0, warm = -1.0, min_nb = -1.0;
    int max_it = -1;
};

// Everything one pressure solve launches with (solve_driver.py:121-162), resolved on the host once.
struct SolveSetup {
    Ctx c;  // by value: overrides stay local to this solve
    int mode;
    float dt, warm, relief;
    float* store;
    const float* learned;
    StopSpec sp;
    double tolh;
};

SolveSetup solve_setup(Handle* H, int mode, double dt_h, int fixed_iters, const SolveOverride& ov) {
    SolveSetup S;
    S.c = H->c;
    const DfsphParams& P = H->P;
    bool div = mode == 0;
    S.mode = mode;
    S.dt = float(dt_h);
    S.store = div ? S.c.kv : S.c.kd;
    S.tolh = div ? P.max_divergence_error : P.max_density_error;
    if (ov.tol >= 0.0) S.tolh = ov.tol;
    if (ov.min_nb >= 0.0) S.c.min_nb = float(ov.min_nb);
    double err_factor = ov.err_factor >= 0.0 ? ov.err_factor : P.max_error_factor;
    double warm_h = div ? P.warm_start_divergence : P.warm_start_density;
    if (ov.warm >= 0.0) warm_h = ov.warm;
    S.warm = float(warm_h);
    S.relief = 1.0e30f;
    if (!div && P.compression_relief_speed > 0.0)
        S.relief = float(P.compression_relief_speed * dt_h * P.lattice_grad);
    S.learned = div ? H->learned_div : H->learned_den;
    S.sp.tol = float(S.tolh);
    S.sp.tol_max = float(S.tolh * err_factor);
    S.sp.min_it = P.min_iterations;
    S.sp.max_it = ov.max_it > 0 ? ov.max_it : P.max_iterations;
    S.sp.use_scale = div ? 0 : 1;
    if (fixed_iters >= 0) S.sp.min_it = S.sp.max_it = std::max(fixed_iters, 1);
    return S;
}

// tolerance scale, source / warm start, begin_solve and (CG) the initial residual
int enq_solve_head(Handle* H, const SolveSetup& S, hipStream_t st) {
    const Ctx& c = S.c;
    const DfsphParams& P = H->P;
    if (S.mode == 1 && P.rest_jitter_speed > 0.0) {  // _density_tolerance (solve_driver.py:108-119)
        k_head_terms<<<H->n_chunks, 256, 0, st>>>(c);
        k_world<<<c.W, 256, 0, st>>>(c, W_HEAD, StopSpec{});
        k_tol_scale<<<blocks_for(c.W), 256, 0, st>>>(c, S.dt, float(P.rest_jitter_speed),
                                                      float(P.rest_jitter_flow_ratio), float(P.max_density_error),
                                                      float(std::min(P.min_density_error, P.max_density_error)));
    }
    LAUNCH_WT(KID_SOLVE_INIT, k_solve_init, H->n_chunks, 256, c, S.mode, S.dt, S.warm,
              float(P.warm_start_density), S.relief, (const float*)S.store, S.learned);
    k_begin_solve<<<blocks_for(c.W), 256, 0, st>>>(c);
    if (P.pressure_solver == 0) {
        LAUNCH_ACCEL(c, c.kappa, c.kr, VR_KAPPA, 0);
        LAUNCH_APPLY(c, S.mode, S.dt, c.kappa, c.wv, 0, 1, 0, 0);
        k_cg_init<<<H->n_chunks, 256, 0, st>>>(c, S.mode, S.dt);
        k_world<<<c.W, 256, 0, st>>>(c, W_CG_INIT, S.sp);
    }
    LAUNCH_CHECK();
    return DFSPH_OK;
}

// check_every iterations (krylov.py:298-313 / solve_driver.py:175-184); converged worlds' blocks exit at once
int enq_solve_batch(Handle* H, const SolveSetup& S, hipStream_t st, int* flag_out) {
    const Ctx& c = S.c;
    int mode = S.mode;
    float dt = S.dt;
    for (int rep = 0; rep < H->check_every; ++rep) {
        if (H->P.pressure_solver == 0) {
            LAUNCH_ACCEL(c, c.pd, c.pdr, VR_DIR, 0);
            LAUNCH_APPLY(c, mode, dt, c.pd, c.ap, 0, 0, 0, 1);
            k_world<<<c.W, 256, 0, st>>>(c, W_CG_PAP, S.sp);
            k_cg_step<<<H->n_chunks, 256, 0, st>>>(c);
            LAUNCH_ACCEL(c, c.corr, c.corrr, VR_CORR, 1);
            LAUNCH_APPLY(c, mode, dt, c.corr, c.wv, 1, 0, 1, 0);
            k_cg_finish<<<H->n_chunks, 256, 0, st>>>(c, mode, dt);
            k_world<<<c.W, 256, 0, st>>>(c, W_CG_BETA, S.sp);
            k_cg_direction<<<H->n_chunks, 256, 0, st>>>(c, rep == H->check_every - 1 ? flag_out : nullptr);
        } else {
            LAUNCH_WT(KID_ACCEL, k_accel, H->n_chunks, 256, c, c.kappa, c.kr, VR_KAPPA, 0);
            LAUNCH_W(k_jacobi, H->n_chunks, 256, c, mode, dt);
            k_world<<<c.W, 256, 0, st>>>(c, W_JACOBI, S.sp);
        }
    }
    LAUNCH_CHECK();
    return DFSPH_OK;
}

// contact limit, apply pressure, lone friction; stats (solve_driver.py:151-162)
int enq_solve_tail(Handle* H, const SolveSetup& S, hipStream_t st) {
    const Ctx& c = S.c;
    LAUNCH_WT(KID_POST, k_post_solve, H->n_chunks, 256, c, S.mode, S.dt, S.store);
    k_record<<<blocks_for(c.W), 256, 0, st>>>(c, S.mode, float(S.tolh));
    LAUNCH_CHECK();
    return DFSPH_OK;
}

// The CG iterations of a solve as one cooperative launch of k_cg_loop (every block resident: hipLaunchCooperativeKernel
// refuses a grid that does not fit), with the barrier counter cleared by a kernel on the same stream first.
template <bool WIDE>
int launch_cg_loop_w(Handle* H, const SolveSetup& S, hipStream_t st) {
    int& grid = H->loop_grid[WIDE ? 1 : 0];
    if (grid == 0) {
        int dev = 0, per_cu = 0;
        hipDeviceProp_t prop;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipGetDeviceProperties(&prop, dev));
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg_loop<WIDE>, 256, 0));
        grid = (per_cu * prop.multiProcessorCount) / N_XCD * N_XCD;
        if (grid < N_XCD) return fail(DFSPH_EDEVICE, "device solve loop: no co-resident grid");
    }
    k_fill_i32<<<1, 256, 0, st>>>(reinterpret_cast<int*>(S.c.loop_bar), 0, LOOP_WORDS);
    Ctx c = S.c;
    int mode = S.mode;
    float dt = S.dt;
    StopSpec sp = S.sp;
    void* args[] = {&c, &mode, &dt, &sp};
    HIPCHK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_cg_loop<WIDE>), dim3(grid), dim3(256), args, 0,
                                      st));
    return DFSPH_OK;
}
int launch_cg_loop(Handle* H, const SolveSetup& S, hipStream_t st) {
    return S.c.wide ? launch_cg_loop_w<true>(H, S, st) : launch_cg_loop_w<false>(H, S, st);
}

// A captured segment of the step: enqueued once into a hipGraph on the handle's private capture stream, then
// replayed on the caller's stream. Kernel arguments (state pointers, dt, iteration batch) are part of the key.
template <class F>
int run_segment(Handle* H, hipStream_t st, int kind, int mode, float dt, F enqueue) {
    if (!H->use_graphs || H->prof.on) return enqueue(st);
    uint32_t bits;
    std::memcpy(&bits, &dt, 4);
    GraphKey key{kind, mode, H->check_every,
                 H->c.pair_cache | (H->c.sparse_corr << 8) | (H->build_block << 16) | (H->c.packed << 28) |
                     (H->c.use_halo << 29) | (H->device_loop << 30) | int(uint32_t(H->c.learn_gate) << 31),
                 bits, H->c.q, H->c.qd, H->c.f, H->learned_div, H->learned_den,
                 H->mlp_params, H->log_feat, H->log_targets};
    auto it = H->graphs.find(key);
    if (it == H->graphs.end()) {
        if (!H->cap_stream) HIPCHK(hipStreamCreateWithFlags(&H->cap_stream, hipStreamNonBlocking));
        HIPCHK(hipStreamBeginCapture(H->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue(H->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(H->cap_stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(DFSPH_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
        hipGraphExec_t ex = nullptr;
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(DFSPH_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
        it = H->graphs.emplace(key, ex).first;
    }
    HIPCHK(hipGraphLaunch(it->second, st));
    return DFSPH_OK;
}

enum SegKind { SEG_PRE = 0, SEG_HEAD, SEG_BATCH, SEG_TAIL_DIV, SEG_TAIL_DEN, SEG_NONPRESSURE };

// one pressure solve with the device-side convergence loop; `tail_extra` runs after the solve's own tail (in the
// same captured segment)
template <class Extra>
int run_solve(Handle* H, const SolveSetup& S, hipStream_t st, int tail_kind, Extra tail_extra) {
    int rc;
    if ((rc = run_segment(H, st, SEG_HEAD, S.mode, S.dt, [&](hipStream_t s) { return enq_solve_head(H, S, s); })))
        return rc;
    if (H->device_loop && H->P.pressure_solver == 0 && !H->prof.on) {  // no host round trip at all
        if ((rc = launch_cg_loop(H, S, st))) return rc;
        return run_segment(H, st, tail_kind, S.mode, S.dt, [&](hipStream_t s) {
            int r = enq_solve_tail(H, S, s);
            return r ? r : tail_extra(s);
        });
    }
    // a captured batch cannot alternate its slot pointer, so replayed graphs and Jacobi batches copy the count
    bool direct = H->P.pressure_solver == 0 && (!H->use_graphs || H->prof.on);
    int nu = 0;
    for (int batch = 0;; ++batch) {
        int slot = batch & 1;
        int* fl = direct ? H->d_unconv + slot : nullptr;
        if ((rc = run_segment(H, st, SEG_BATCH, S.mode, S.dt,
                              [&](hipStream_t s) { return enq_solve_batch(H, S, s, fl); })))
            return rc;
        if ((rc = post_unconv(H, st, slot, direct))) return rc;
        if (batch == 0) continue;
        if ((rc = wait_unconv(H, slot ^ 1, &nu))) return rc;
        if (nu == 0) break;
        if (batch > 100000) return fail(DFSPH_EDEVICE, "pressure solve did not terminate");
    }
    return run_segment(H, st, tail_kind, S.mode, S.dt, [&](hipStream_t s) {
        int r = enq_solve_tail(H, S, s);
        return r ? r : tail_extra(s);
    });
}

int phase_solve(Handle* H, int mode, double dt_h, int fixed_iters, hipStream_t st, const SolveOverride& ov = {}) {
    if (H->c.n == 0) return DFSPH_OK;
    SolveSetup S = solve_setup(H, mode, dt_h, fixed_iters, ov);
    bool graphs = H->use_graphs;
    H->use_graphs = false;  // phase entry points and prime run eagerly
    int rc = run_solve(H, S, st, mode == 0 ? SEG_TAIL_DIV : SEG_TAIL_DEN, [](hipStream_t) { return 0; });
    H->use_graphs = graphs;
    return rc;
}

// csf.py:311-332 CSFTension.apply: one stage of the non-pressure update
int run_csf(Handle* H, float dt, hipStream_t st) {
    Ctx& c = H->c;
    if (!c.csf || c.n == 0) return DFSPH_OK;
    const int nb = blocks_for(c.n);
    LAUNCH_W(k_csf_grad, nb, 256, c);
    // NORMAL_PASSES + 1 = 3 Shepard passes: grad -> g1 -> g2 -> g1
    LAUNCH_W(k_csf_shepard_vec, nb, 256, c, c.t_grad, c.t_g1);
    LAUNCH_W(k_csf_shepard_vec, nb, 256, c, c.t_g1, c.t_g2);
    LAUNCH_W(k_csf_shepard_vec, nb, 256, c, c.t_g2, c.t_g1);
    k_csf_normal<<<nb, 256, 0, st>>>(c, c.t_g1);
    LAUNCH_W(k_csf_curvature, nb, 256, c);
    // CURVATURE_PASSES = 3: kraw -> ktmp -> kraw -> ktmp
    LAUNCH_W(k_csf_shepard_surf, nb, 256, c, c.t_kraw, c.t_ktmp);
    LAUNCH_W(k_csf_shepard_surf, nb, 256, c, c.t_ktmp, c.t_kraw);
    LAUNCH_W(k_csf_shepard_surf, nb, 256, c, c.t_kraw, c.t_ktmp);
    k_csf_apply<<<nb, 256, 0, st>>>(c, c.t_ktmp, dt);
    LAUNCH_W(k_csf_akinci_normals, nb, 256, c);
    LAUNCH_W(k_csf_pairs, nb, 256, c, dt);
    LAUNCH_CHECK();
    return DFSPH_OK;
}

int phase_nonpressure(Handle* H, double dt, hipStream_t st) {
    Ctx& c = H->c;
    if (c.n == 0) return DFSPH_OK;
    if (c.xsph > 0.0f) k_pack_vrho<<<blocks_for(c.n), 256, 0, st>>>(c);
    LAUNCH_WT(KID_NONPRESSURE, k_nonpressure, H->n_chunks, 256, c, float(dt));
    LAUNCH_CHECK();
    return run_csf(H, float(dt), st);  // solver.py:304-305: surface tension after the non-pressure update
}

// viscosity.py:257-283 ImplicitViscosity.solve: v* = the velocities after the non-pressure update; CG per world with
// the same lookahead loop as the pressure solves; the solution replaces the liquid's velocities
int run_visc(Handle* H, double dt_h, hipStream_t st, bool predict) {
    Ctx& c = H->c;
    if (c.n == 0 || !H->P.viscosity_implicit) return DFSPH_OK;
    const float dt = float(dt_h);
    StopSpec sp{float(H->P.visc_tol), 0.0f, 0, H->P.visc_max_iterations, 0};
    const int64_t n3 = 3 * int64_t(c.n);
    const bool wall = c.wall_coeff > 0.0f;
    if (wall) {  // solver.py:310-324 _wall_terms: with predict, the walls' velocities carry the pressure guess
        if (predict) {
            k_visc_guess<<<blocks_for(c.n), 256, 0, st>>>(c, dt);
            LAUNCH_W(k_visc_raw, blocks_for(c.n), 256, c, dt);
            LAUNCH_W(k_visc_shift, blocks_for(c.n), 256, c);
        }
        k_visc_wall_terms<<<blocks_for(c.n), 256, 0, st>>>(c, predict ? 1 : 0);
    }
    k_copy_f32<<<blocks_for(n3), 256, 0, st>>>(c.vb, c.qd, n3);
    k_copy_f32<<<blocks_for(n3), 256, 0, st>>>(c.vx, c.qd, n3);
    k_fill_i32<<<blocks_for(c.W), 256, 0, st>>>(c.v_conv, 0, c.W);
    LAUNCH_W(k_visc_apply, H->n_chunks, 256, c, dt, c.vx, c.vap);
    LAUNCH_W(k_visc_init, H->n_chunks, 256, c, dt);
    k_fill_i32<<<1, 256, 0, st>>>(c.n_unconv, 0, 1);
    k_world<<<c.W, 256, 0, st>>>(c, W_VISC_BEGIN, sp);
    int rc, nu = 0;
    for (int batch = 0;; ++batch) {
        int slot = batch & 1;
        LAUNCH_W(k_visc_apply, H->n_chunks, 256, c, dt, c.vp, c.vap);
        k_world<<<c.W, 256, 0, st>>>(c, W_VISC_ALPHA, sp);
        k_visc_update<<<H->n_chunks, 256, 0, st>>>(c);
        k_world<<<c.W, 256, 0, st>>>(c, W_VISC_CHECK, sp);
        k_visc_direction<<<H->n_chunks, 256, 0, st>>>(c, H->d_unconv + slot);
        LAUNCH_CHECK();
        if ((rc = post_unconv(H, st, slot, true))) return rc;
        if (batch == 0) continue;
        if ((rc = wait_unconv(H, slot ^ 1, &nu))) return rc;
        if (nu == 0) break;
        if (batch > 100000) return fail(DFSPH_EDEVICE, "viscosity solve did not terminate");
    }
    k_visc_store<<<blocks_for(c.n), 256, 0, st>>>(c);
    if (wall) k_visc_wall_reaction<<<blocks_for(c.n), 256, 0, st>>>(c, dt, predict ? 1 : 0);
    LAUNCH_CHECK();
    return DFSPH_OK;
}

int phase_advect(Handle* H, double dt, hipStream_t st) {
    Ctx& c = H->c;
    if (c.n == 0) return DFSPH_OK;
    k_fill_f32<<<blocks_for(c.W), 256, 0, st>>>(c.st_max_speed, 0.0f, c.W);
    LAUNCH_T(KID_ADVECT, k_advect_project, H->n_chunks, 256, c, float(dt));
    LAUNCH_CHECK();
    return DFSPH_OK;
}

// particle_walls.py:246-259 begin (work zeroed) and accumulate: the walls' share of the step's slot impulses
int phase_walls(Handle* H, hipStream_t st) {
    Ctx& c = H->c;
    if (c.n == 0 || c.NW == 0) return DFSPH_OK;
    k_fill_i32<<<blocks_for(2 * int64_t(c.W)), 256, 0, st>>>(reinterpret_cast<int*>(c.pw_work), 0, 2 * int64_t(c.W));
    k_pw_accumulate<<<blocks_for(c.n), 256, 0, st>>>(c);
    LAUNCH_CHECK();
    return DFSPH_OK;
}

// runtime.py:318-347 EmitterManager.step: sinks and the escape box, then the box emitters; the speed statistics of
// the step's end state come from these passes (c.speed_later)
int phase_emitters(Handle* H, double dt, hipStream_t st) {
    Ctx& c = H->c;
    if (c.n == 0 || !H->emit_active) return DFSPH_OK;
    k_retire<<<H->n_chunks, 256, 0, st>>>(c, H->em);
    if (H->em.E > 0) k_emit<<<H->em.E, 256, 0, st>>>(c, H->em, float(dt));
    LAUNCH_CHECK();
    return DFSPH_OK;
}

void drop_graphs(Handle* H) {
    for (auto& kv : H->graphs) (void)hipGraphExecDestroy(kv.second);
    H->graphs.clear();
}

void bind_state(Handle* H, const float* q, const float* qd, const float* f) {
    H->c.q = const_cast<float*>(q);
    H->c.qd = const_cast<float*>(qd);
    H->c.f = f;
}

// the learned warm start feeds the solves DFSPH_OPT_LEARNED_SOLVES selects
void apply_learned(Handle* H) {
    bool on = H->mlp_params != nullptr;
    H->learned_div = on && (H->learned_solves & 1) ? H->c.warm_k : nullptr;
    H->learned_den = on && (H->learned_solves & 2) ? H->c.warm_k + H->c.n : nullptr;
}

}  // namespace

// =============================================================================================== C-ABI
extern "C" {

int32_t dfsph_abi_version(void) { return DFSPH_ABI_VERSION; }
const char* dfsph_last_error(void) { return g_err.c_str(); }

int32_t dfsph_create(const DfsphParams* P, const DfsphLayout* L, const DfsphModel* M, void** out) {
    if (!P || !L || !M || !out) return fail(DFSPH_EINVAL, "dfsph_create: null argument");
    if (L->n < 0 || L->world_count < 1) return fail(DFSPH_EINVAL, "dfsph_create: need n >= 0 and world_count >= 1");
    if (L->n >= (int64_t(1) << 28))  // the solve gathers' 32-bit buffer offsets (4 B entries of NB_U rows)
        return fail(DFSPH_EINVAL, "dfsph_create: at most 2^28 - 1 particles per handle");
    if (P->neighbor_capacity < 8 || P->neighbor_capacity > 64)
        return fail(DFSPH_EINVAL, "neighbor_capacity must be in [8, 64]");
    if (!(P->h > 0.0)) return fail(DFSPH_EINVAL, "h must be > 0");
    if (P->skin > 0.0 && L->wall_particle_count > 0)  // neighbors.py:249-254
        return fail(DFSPH_EINVAL, "SolverDFSPH: particle walls need the grid rebuilt every step (reneighbor_skin off)");
    if (!(P->skin >= 0.0) || P->skin > P->h) return fail(DFSPH_EINVAL, "reneighbor_skin must be in [0, 1] (x h)");
    if (!(P->boundary_viscosity >= 0.0) || !(P->visc_wall_coeff >= 0.0))
        return fail(DFSPH_EINVAL, "boundary_viscosity must be >= 0");
    if (P->boundary_viscosity > 0.0 && !P->viscosity_implicit)  // solver.py:164-165
        return fail(DFSPH_EINVAL, "boundary_viscosity needs the implicit (weiler2018) viscosity");
    const int W = L->world_count, n = L->n, S = L->shape_count;
    if (L->world_start[0] != 0 || L->world_start[W] != n)
        return fail(DFSPH_EINVAL, "world_start must cover every particle (world_start[0] = 0, world_start[W] = n)");
    for (int w = 0; w < W; ++w)
        if (L->world_start[w + 1] < L->world_start[w]) return fail(DFSPH_EINVAL, "world_start must be ascending");
    for (int s = 0; s < S; ++s) {
        if (L->shape_body[s] >= 0)
            return fail(DFSPH_EINVAL, "shapes attached to bodies are not supported (static shapes only)");
        int g = L->shape_geo[s];
        if (g != GEO_BOX && g != GEO_SPHERE && g != GEO_PLANE)
            return fail(DFSPH_EINVAL, "unsupported shape geometry (box, sphere and infinite plane only)");
        if (g == GEO_PLANE && (L->shape_scale[3 * s] > 0.0f && L->shape_scale[3 * s + 1] > 0.0f))
            return fail(DFSPH_EINVAL, "finite planes are not supported");
    }
    Handle* H = new Handle();
    H->P = *P;
    Ctx& c = H->c;
    std::memset(&c, 0, sizeof(Ctx));
    c.n = n;
    c.W = W;
    c.sparse_corr = 1;
    c.S = S;
    c.NW = L->wall_particle_count;
    c.one_way = P->one_way != 0;
    c.mu = M->viscosity;
    c.visc_d = float(P->visc_d);
    c.wall_coeff = P->viscosity_implicit ? float(P->visc_wall_coeff) : 0.0f;
    c.mu_b = float(P->boundary_viscosity);
    c.csf = P->csf != 0;
    c.csf_cs = float(P->csf_c_surface);
    c.csf_inv_i = float(P->csf_inv_i);
    c.csf_lo = float(P->csf_band_lo);
    c.csf_hi = float(P->csf_band_hi);
    c.csf_kcoh = float(P->csf_k_cohesion);
    c.cap = P->neighbor_capacity;
    c.kind = P->kernel;
    c.friction = P->wall_friction;
    c.curvature = P->curvature;
    c.two_sided = P->divergence_two_sided;
    c.h = float(P->h);
    c.cell = float((P->h + std::max(P->skin, 0.0)) * (1.0 + 1.0e-5));  // a skin widens the cells to h + skin
    c.skin = float(std::max(P->skin, 0.0));
    c.r = float(P->particle_radius);
    c.n_full = float(P->n_full);
    c.tot_min = float(P->tot_min);
    c.min_nb = float(P->min_neighbors_divergence);
    c.omega = float(P->jacobi_omega);
    c.warm_v = float(P->warm_start_divergence);
    c.warm_d = float(P->warm_start_density);
    c.xsph = float(P->xsph_c);
    c.v_sep = float(P->contact_separation_speed);
    c.offset = float(P->wall_offset);
    c.amplitude = float(P->wall_amplitude);
    c.taper0 = float(P->taper0);
    c.taper1 = float(P->taper1);
    c.reach = float(P->reach);
    c.eps = float(P->project_eps);
    c.sweep_len = float(P->sweep_len);
    {
        double gmax = 0.0;  // |g| of the first world with gravity sets the kappa scale (9.81 when none)
        std::vector<float> gh(3 * size_t(W));
        if (hipMemcpy(gh.data(), M->gravity, gh.size() * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess)
            for (int w = 0; w < W && gmax == 0.0; ++w)
                gmax = std::sqrt(double(gh[3 * w]) * gh[3 * w] + double(gh[3 * w + 1]) * gh[3 * w + 1] +
                                 double(gh[3 * w + 2]) * gh[3 * w + 2]);
        c.kscale = float((gmax > 0.0 ? gmax : 9.81) * P->h);
        c.lat_diag = float(P->lattice_grad * P->lattice_grad);
    }
    c.mass = M->mass;
    c.rho0 = M->rest_density;
    c.flags = M->flags;
    c.phase = M->phase;
    c.world = M->world;
    c.gravity = M->gravity;
    c.wvol = M->wall_volume;
    int maxw = 0;
    for (int w = 0; w < W; ++w) maxw = std::max(maxw, L->world_start[w + 1] - L->world_start[w]);
    c.wide = maxw > 65536;
    int rc = 0;
#define TRY(x)             \
    do {                   \
        rc = (x);          \
        if (rc) goto fail; \
    } while (0)
    {
        // chunk table
        std::vector<int> cb, ce, cw, wc{0};
        for (int w = 0; w < W; ++w) {
            for (int b = L->world_start[w]; b < L->world_start[w + 1]; b += CHUNK) {
                cb.push_back(b);
                ce.push_back(std::min(b + CHUNK, L->world_start[w + 1]));
                cw.push_back(w);
            }
            wc.push_back(int(cb.size()));
        }
        H->n_chunks = int(cb.size());
        c.n_chunks = H->n_chunks;
        TRY(dupload(H, (int**)&c.ch_begin, cb.data(), int64_t(cb.size())));
        TRY(dupload(H, (int**)&c.ch_end, ce.data(), int64_t(ce.size())));
        TRY(dupload(H, (int**)&c.ch_world, cw.data(), int64_t(cw.size())));
        TRY(dupload(H, (int**)&c.w_chunk, wc.data(), int64_t(wc.size())));
        TRY(dalloc(H, &c.ch_live, int64_t(cb.size())));
        std::vector<int> ws(L->world_start, L->world_start + W + 1);
        TRY(dupload(H, (int**)&c.ws, ws.data(), W + 1));
        // grid
        std::vector<float> org(L->grid_origin, L->grid_origin + 3 * W);
        std::vector<int> dims(3 * W), base(2 * W + 1);
        long long tot = 0;
        for (int k = 0; k < 3 * W; ++k) dims[k] = std::max(1, L->grid_dims[k]);
        for (int g = 0; g < 2 * W; ++g) {
            int w = g % W;
            base[g] = int(tot);
            tot += (long long)dims[3 * w] * dims[3 * w + 1] * dims[3 * w + 2];
        }
        if (tot > (1LL << 30)) {
            rc = fail(DFSPH_ECAPACITY, "cell grid too large (> 2^30 cells)");
            goto fail;
        }
        base[2 * W] = int(tot);
        c.n_cells = int(tot);
        c.wall_first = base[W];  // first cell of the wall grids
        TRY(dupload(H, (float**)&c.g_org, org.data(), 3 * W));
        TRY(dupload(H, (int**)&c.g_dims, dims.data(), 3 * W));
        TRY(dupload(H, (int**)&c.cell_base, base.data(), 2 * W + 1));
        TRY(dalloc(H, &c.cell_count, int64_t(c.n_cells) + 1));
        TRY(dalloc(H, &c.cell_start, int64_t(c.n_cells) + 1));
        if (c.NW > 0) TRY(dalloc(H, &c.wall_near, int64_t(c.n_cells) - c.wall_first));
        {  // compact wall index (ascending ids), the exchange window and the walls' work
            std::vector<float> wv(std::max(n, 1), 0.0f);
            if (n > 0) HIPCHK(hipMemcpy(wv.data(), M->wall_volume, sizeof(float) * n, hipMemcpyDeviceToHost));
            std::vector<int> cidx(std::max(n, 1), -1);
            int nw = 0;
            for (int i = 0; i < n; ++i)
                if (wv[i] > 0.0f) cidx[i] = nw++;
            H->n_wall = nw;
            TRY(dupload(H, (int**)&c.wcidx, cidx.data(), std::max(n, 1)));
            TRY(dalloc(H, &c.pw_exch, 3 * int64_t(std::max(nw, 1))));
            TRY(dalloc(H, &c.pw_work, W));
            TRY(dalloc(H, &c.pw_col, nw > 0 ? n : 1));
            TRY(dalloc(H, &c.wr_n, nw > 0 ? n : 1));
            TRY(dalloc(H, &c.wr_id, nw > 0 ? int64_t(PW_ROW) * n : 1));
            TRY(dalloc(H, &c.wr_w, nw > 0 ? int64_t(PW_ROW) * n : 1));
        }
        H->n_scan_blocks = blocks_for(c.n_cells, SCAN_B);
        TRY(dalloc(H, &c.scan_part, H->n_scan_blocks));
        TRY(dalloc(H, &c.items, n));
        TRY(dalloc(H, &c.items_u, n));
        TRY(dalloc(H, &c.q4s_u, n));
        TRY(dalloc(H, &c.vrho, n));
        TRY(dalloc(H, &c.key, n));
        TRY(dalloc(H, &c.arrive, n));
        // shapes
        std::vector<int> ss(L->shape_start, L->shape_start + W + 1);
        TRY(dupload(H, (int**)&c.sh_start, ss.data(), W + 1));
        TRY(dupload(H, (int**)&c.sh_ids, L->shape_ids, ss[W]));
        TRY(dupload(H, (int**)&c.sh_geo, L->shape_geo, S));
        TRY(dupload(H, (float**)&c.sh_scale, L->shape_scale, 3 * S));
        TRY(dupload(H, (float**)&c.sh_x, L->shape_xform, 7 * S));
        TRY(dupload(H, (float**)&c.sh_radius, L->shape_radius, S));
        TRY(dupload(H, (float**)&c.sh_mu, L->shape_mu, S));
        {  // contact angles (csf.py:281-296); none given: every contact line free
            std::vector<float> cs(std::max(S, 1), 0.0f), sn(std::max(S, 1), -1.0f);
            if (L->shape_cos && L->shape_sin)
                for (int s = 0; s < S; ++s) {
                    cs[s] = L->shape_cos[s];
                    sn[s] = L->shape_sin[s];
                }
            TRY(dupload(H, &c.sh_cos, cs.data(), std::max(S, 1)));
            TRY(dupload(H, &c.sh_sin, sn.data(), std::max(S, 1)));
        }
        // lists
        H->nbr_bytes = int64_t(c.cap) * std::max(n, 1) * 2;
        H->nbr_hi_bytes = c.wide ? int64_t(c.cap) * std::max(n, 1) : 1;
        // ROW_PAD zero rows past the capacity: a neighbour batch reads them unconditionally (nbr_batch), and the
        // staged loops prefetch the next batch's entries (for_slots_g)
        const int64_t rows = int64_t(c.cap) + ROW_PAD;
        const int64_t quad_rows = (rows + 3) / 4 * 4;  // the neighbour rows are stored in quads (row_at)
        TRY(dalloc(H, &c.nbr, quad_rows * std::max(n, 1)));
        TRY(dalloc(H, &c.nbr_hi, c.wide ? quad_rows * std::max(n, 1) : 4));
        TRY(dalloc(H, &c.pgrad, rows * std::max(n, 1)));
        TRY(dalloc(H, &c.h_slot, rows * std::max(n, 1)));
        TRY(dalloc(H, &c.h_ids, int64_t(std::max(H->n_chunks, 1)) * HALO_MAX));
        TRY(dalloc(H, &c.h_count, std::max(H->n_chunks, 1)));
        // no halo until the first list build with the option on: every chunk gathers directly
        k_fill_i32<<<blocks_for(std::max(H->n_chunks, 1)), 256>>>(c.h_count, -1, std::max(H->n_chunks, 1));
        LAUNCH_CHECK();
        {  // one mass per world (bitwise) over its SPH particles: the packed acceleration gathers may stand in for it
            std::vector<float> hm(std::max(n, 1));
            std::vector<int8_t> hp(std::max(n, 1));
            if (n > 0) {
                HIPCHK(hipMemcpy(hm.data(), M->mass, sizeof(float) * size_t(n), hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(hp.data(), M->phase, size_t(n), hipMemcpyDeviceToHost));
            }
            std::vector<float> wm(std::max(W, 1), 0.0f);
            bool ok = true;
            for (int w = 0; w < W && ok; ++w) {
                bool seen = false;
                for (int i = L->world_start[w]; i < L->world_start[w + 1]; ++i) {
                    if (hp[i] < 0) continue;
                    uint32_t a, b;
                    std::memcpy(&a, &hm[i], 4);
                    std::memcpy(&b, &wm[w], 4);
                    if (!seen) {
                        wm[w] = hm[i];
                        seen = true;
                    } else if (a != b) {
                        ok = false;
                        break;
                    }
                }
            }
            H->packed_ok = ok ? 1 : 0;
            float* wm_dev = nullptr;
            TRY(dupload(H, &wm_dev, wm.data(), std::max(W, 1)));
            c.w_mass = wm_dev;
            if (H->packed_ok) {
                TRY(dalloc(H, &c.qk, 3 * int64_t(std::max(n, 1))));
                TRY(dalloc(H, &c.qa, n));
                TRY(dalloc(H, &c.a2, n));
            }
            c.packed = H->packed_ok;
        }
        if (P->viscosity_implicit) {  // viscosity.py:213-234
            for (float** p : {&c.vb, &c.vx, &c.vr, &c.vz, &c.vp, &c.vap}) TRY(dalloc(H, p, 3 * int64_t(n)));
            TRY(dalloc(H, &c.vminv, 9 * int64_t(n)));
            for (int** p : {&c.v_conv, &c.v_iters}) TRY(dalloc(H, p, W));
            for (double** p : {&c.v_rz, &c.v_bb, &c.v_step}) TRY(dalloc(H, p, W));
            if (!M->viscosity) {
                rc = fail(DFSPH_EINVAL, "the implicit viscosity needs the model's sph:viscosity (DfsphModel.viscosity)");
                goto fail;
            }
            if (c.wall_coeff > 0.0f) {  // wall_viscosity.py: per-particle wall block, drag and guess shift
                TRY(dalloc(H, &c.wall_m, 9 * int64_t(n)));
                for (float** p : {&c.wall_w, &c.vshift, &c.vraw}) TRY(dalloc(H, p, 3 * int64_t(n)));
                TRY(dalloc(H, &c.vgk, int64_t(n)));
            }
        }
        if (c.csf) {  // csf.py:298-309: the per-step work arrays
            if (!M->surface_tension || !M->csf_gamma) {
                rc = fail(DFSPH_EINVAL, "surface_tension 'csf' needs DfsphModel.surface_tension and csf_gamma");
                goto fail;
            }
            c.sigma = M->surface_tension;
            c.gamma = M->csf_gamma;
            for (float** p : {&c.t_mcount, &c.t_kraw, &c.t_ktmp, &c.t_kappa}) TRY(dalloc(H, p, n));
            for (float** p : {&c.t_grad, &c.t_g1, &c.t_g2, &c.t_nhat, &c.t_nwall, &c.t_anorm})
                TRY(dalloc(H, p, 3 * int64_t(n)));
            for (int** p : {&c.t_surf, &c.t_wslot}) TRY(dalloc(H, p, n));
        }
        if (c.skin > 0.0f) {  // Verlet-skin state (neighbors_skin.py:162-177)
            c.cap_s = P->skin_rows > 0 ? std::min(P->skin_rows, 64) : 64;
            TRY(dalloc(H, &c.sup, int64_t(c.cap_s) * std::max(n, 1)));
            TRY(dalloc(H, &c.sup_hi, c.wide ? int64_t(c.cap_s) * std::max(n, 1) : 1));
            TRY(dalloc(H, &c.sup_count, n));
            TRY(dalloc(H, &c.q_ref, n));
            std::vector<int> never(std::max(n, 1), INT_MIN);  // never built: every world builds first
            TRY(dupload(H, &c.g_ref, never.data(), std::max(n, 1)));
            TRY(dalloc(H, &c.skin_flag, W));
            TRY(dalloc(H, &c.skin_force, W));
        }
        TRY(dalloc(H, &c.count, n));
        TRY(dalloc(H, &c.group, n));
        // slots
        TRY(dalloc(H, &c.s_shape, 4 * int64_t(n)));
        TRY(dalloc(H, &c.s_psi, 4 * int64_t(n)));
        TRY(dalloc(H, &c.s_phi, 4 * int64_t(n)));
        for (float** p : {&c.s_gpsi, &c.s_xb, &c.s_vb, &c.s_xbl, &c.s_nbl, &c.s_imp}) TRY(dalloc(H, p, 12 * int64_t(n)));
        for (float** p : {&c.ratio, &c.alpha, &c.neff, &c.active, &c.reg, &c.src, &c.kappa, &c.err, &c.lower, &c.kv,
                          &c.kd, &c.pd, &c.z, &c.wv, &c.ap, &c.corr, &c.kr, &c.pdr, &c.corrr})
            TRY(dalloc(H, p, n));
        TRY(dalloc(H, &c.acc4, n));
        TRY(dalloc(H, &c.touched, n));
        TRY(dalloc(H, &c.q4, n));
        TRY(dalloc(H, &c.q4s, int64_t(n) + BUILD_U));  // BUILD_U records of padding: the build loads batches unclamped
        TRY(dalloc(H, &c.feat, 16 * int64_t(n)));
        TRY(dalloc(H, &c.mlp_out, 2 * int64_t(n)));
        TRY(dalloc(H, &c.warm_k, 2 * int64_t(n)));
        TRY(dalloc(H, &c.learn_on, std::max(W, 1)));
        TRY(dalloc(H, &H->learn_tiles, (int64_t(n) + 63) / 64));
        TRY(dalloc(H, &c.vtmp, 3 * int64_t(n)));
        TRY(dalloc(H, &c.act, n));
        for (int** p : {&c.converged, &c.iters, &c.restart, &c.st_overflow, &c.st_projected, &c.st_swept,
                        &c.st_nonfinite, &c.it_den, &c.it_div})
            TRY(dalloc(H, p, W));
        for (float** p : {&c.err_mean, &c.err_max, &c.tol_scale, &c.st_max_speed, &c.st_den_err, &c.st_div_err,
                          &c.st_den_err_max, &c.st_den_tol})
            TRY(dalloc(H, p, W));
        for (double** p : {&c.n_active, &c.err_sum, &c.rz, &c.stepb}) TRY(dalloc(H, p, W));
        for (double** p : {&c.sums, &c.flow_z, &c.flow_xy, &c.head}) TRY(dalloc(H, p, 2 * W));
        TRY(dalloc(H, &c.meter, 3 * int64_t(std::max(S, 1))));
        {  // ledgers (runtime.py:93-107) and an empty sink table
            EmitCtx& e = H->em;
            for (long long** p : {&e.mass0, &e.mass_in, &e.mass_out, &e.mass_act, &e.flux_e, &e.flux_g})
                TRY(dalloc(H, p, W));
            TRY(dalloc(H, &e.escaped, W));
            TRY(dalloc(H, &e.emit_overflow, W));
            int* sink_start = nullptr;
            TRY(dalloc(H, &sink_start, W + 1));
            e.sink_start = sink_start;
        }
        TRY(dalloc(H, &c.part, 3 * int64_t(std::max(H->n_chunks, 1))));
        TRY(dalloc(H, &c.n_unconv, 1));
        TRY(dalloc(H, &c.loop_bar, LOOP_WORDS));
        TRY(dalloc(H, &c.loop_fault, 1));
        TRY(dalloc(H, &H->qd_saved, 3 * int64_t(n)));
        {
            std::vector<float> ones(W, 1.0f);
            HIPCHK(hipMemcpy(c.tol_scale, ones.data(), W * sizeof(float), hipMemcpyHostToDevice));
        }
        void* pinned = nullptr;
        if (hipHostMalloc(&pinned, 2 * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
            rc = fail(DFSPH_EDEVICE, "hipHostMalloc failed");
            goto fail;
        }
        H->h_unconv = static_cast<int*>(pinned);
        H->h_unconv[0] = H->h_unconv[1] = 1;
        void* dview = nullptr;
        if (hipHostGetDevicePointer(&dview, pinned, 0) != hipSuccess) {
            rc = fail(DFSPH_EDEVICE, "hipHostGetDevicePointer failed");
            goto fail;
        }
        H->d_unconv = static_cast<int*>(dview);
        for (int k = 0; k < 2; ++k)
            if (hipEventCreateWithFlags(&H->ev_unconv[k], hipEventDisableTiming) != hipSuccess) {
                rc = fail(DFSPH_EDEVICE, "hipEventCreate failed");
                goto fail;
            }
        // runtime.py:132-134 init_mass: the build-time active mass is the ledger's initial mass
        if (n > 0) {
            k_active_mass<<<H->n_chunks, 256, 0, nullptr>>>(c, H->em.mass0);
            if (hipDeviceSynchronize() != hipSuccess) {
                rc = fail(DFSPH_EDEVICE, "initial mass failed");
                goto fail;
            }
        }
    }
    *out = H;
    return DFSPH_OK;
fail:
    for (void* p : H->owned) (void)hipFree(p);
    delete H;
    return rc;
#undef TRY
}

int32_t dfsph_destroy(void* h) {
    if (!h) return DFSPH_OK;
    Handle* H = as_handle(h);
    (void)hipDeviceSynchronize();
    (void)prof_harvest(H, true);
    for (hipEvent_t e : H->prof.pool) (void)hipEventDestroy(e);
    for (auto& kv : H->graphs) (void)hipGraphExecDestroy(kv.second);
    if (H->cap_stream) (void)hipStreamDestroy(H->cap_stream);
    for (void* p : H->owned) (void)hipFree(p);
    for (void* p : H->emit_owned) (void)hipFree(p);
    if (H->h_unconv) (void)hipHostFree(H->h_unconv);
    for (hipEvent_t e : H->ev_unconv)
        if (e) (void)hipEventDestroy(e);
    delete H;
    return DFSPH_OK;
}

int32_t dfsph_build_neighbors(void* h, const float* q, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    Handle* H = as_handle(h);
    bind_state(H, q, H->c.qd, nullptr);
    return phase_neighbors(H, static_cast<hipStream_t>(stream));
}

int32_t dfsph_boundary_slots(void* h, const float* q, const float* qd, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    Handle* H = as_handle(h);
    bind_state(H, q, qd, nullptr);
    return phase_slots(H, static_cast<hipStream_t>(stream));
}

int32_t dfsph_density_alpha(void* h, const float* q, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    Handle* H = as_handle(h);
    bind_state(H, q, H->c.qd, nullptr);
    return phase_density(H, static_cast<hipStream_t>(stream));
}

int32_t dfsph_solve(void* h, int32_t mode, float* q, float* qd, double dt, int32_t fixed_iters, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    if (!(dt > 0.0)) return fail(DFSPH_EINVAL, "dt must be > 0");
    if (mode != 0 && mode != 1) return fail(DFSPH_EINVAL, "mode must be 0 (divergence) or 1 (density)");
    Handle* H = as_handle(h);
    bind_state(H, q, qd, nullptr);
    return phase_solve(H, mode, dt, fixed_iters, static_cast<hipStream_t>(stream));
}

int32_t dfsph_nonpressure(void* h, const float* q, float* qd, const float* f, double dt, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    if (!(dt > 0.0)) return fail(DFSPH_EINVAL, "dt must be > 0");
    Handle* H = as_handle(h);
    bind_state(H, q, qd, f);
    hipStream_t st = static_cast<hipStream_t>(stream);
    int rc = phase_nonpressure(H, dt, st);
    return rc ? rc : run_visc(H, dt, st, true);  // solver.py:295-308
}

int32_t dfsph_advect_project(void* h, float* q, float* qd, double dt, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    if (!(dt > 0.0)) return fail(DFSPH_EINVAL, "dt must be > 0");
    Handle* H = as_handle(h);
    bind_state(H, q, qd, nullptr);
    return phase_advect(H, dt, static_cast<hipStream_t>(stream));
}

int32_t dfsph_flow_speed(void* h, const float* qd, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    Handle* H = as_handle(h);
    bind_state(H, H->c.q, qd, nullptr);
    return phase_flow(H, static_cast<hipStream_t>(stream));
}

// solver.py:218-268
int32_t dfsph_step(void* h, float* q, float* qd, const float* f, double dt, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    if (!(dt > 0.0)) return fail(DFSPH_EINVAL, "SolverDFSPH.step: dt must be > 0");
    Handle* H = as_handle(h);
    hipStream_t st = static_cast<hipStream_t>(stream);
    bind_state(H, q, qd, f);
    if (H->c.n == 0) return DFSPH_OK;
    if (H->prof.on && prof_harvest(H, false)) return DFSPH_EDEVICE;
    float dtf = float(dt);
    const bool learned = (H->mlp_params && H->learned_solves != 0) || H->log_feat;
    // the per-world gate applies to the network's own start (not to logging or to given kappa fields); it holds for
    // this step's launches only
    const bool mlp_start = H->mlp_params && (H->learned_div == H->c.warm_k || H->learned_den == H->c.warm_k + H->c.n);
    H->c.learn_gate = (learned && mlp_start && H->learned_gate && !H->log_feat) ? 1 : 0;
    int rc = run_segment(H, st, SEG_PRE, 0, dtf, [&](hipStream_t s) {
        int r;
        if ((r = phase_neighbors(H, s))) return r;
        if ((r = phase_slots(H, s))) return r;
        if (H->c.learn_gate) {
            k_learn_flags<<<blocks_for(H->c.W), 256, 0, s>>>(H->c, H->learned_solves, H->P.min_iterations);
            k_learn_tiles<<<blocks_for(H->c.n), 256, 0, s>>>(H->c, H->learn_tiles);
        }
        if ((r = phase_density(H, s, learned ? dtf : 0.0f))) return r;
        if ((r = phase_flow(H, s))) return r;
        if (learned) {  // learned warm start: features (written by the density pass) -> MLP -> initial kappa
            Ctx& c = H->c;
            if (H->mlp_params) {
                dfsph_launch_mlp_forward(H->mlp_params, c.feat, c.n, c.mlp_out, c.learn_gate ? H->learn_tiles : nullptr, s);
                k_warm_from_mlp<<<H->n_chunks, 256, 0, s>>>(c, dtf);
            }
            if (H->log_feat)
                k_copy_f32<<<blocks_for(16 * int64_t(c.n)), 256, 0, s>>>(H->log_feat, c.feat, 16 * int64_t(c.n));
            LAUNCH_CHECK();
        }
        return 0;
    });
    if (rc) return rc;
    if (H->P.divergence_solve) {
        SolveSetup S0 = solve_setup(H, 0, dt, -1, SolveOverride{});
        rc = run_solve(H, S0, st, SEG_TAIL_DIV, [&](hipStream_t s) { return phase_nonpressure(H, dt, s); });
    } else {
        rc = run_segment(H, st, SEG_NONPRESSURE, 0, dtf, [&](hipStream_t s) { return phase_nonpressure(H, dt, s); });
    }
    if (rc) return rc;
    if ((rc = run_visc(H, dt, st, true))) return rc;  // solver.py:306-308: part of the non-pressure update
    SolveSetup S1 = solve_setup(H, 1, dt, -1, SolveOverride{});
    rc = run_solve(H, S1, st, SEG_TAIL_DEN, [&](hipStream_t s) {
        if (H->log_targets) k_targets<<<blocks_for(H->c.n), 256, 0, s>>>(H->c, dtf, H->log_targets);
        int r = phase_walls(H, s);
        if (r == 0) r = phase_advect(H, dt, s);
        return r ? r : phase_emitters(H, dt, s);
    });
    H->c.learn_gate = 0;
    return rc;
}

// capture the step's segments in hipGraphs (default on; off runs every launch eagerly)
int32_t dfsph_set_graphs(void* h, int32_t on) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    as_handle(h)->use_graphs = on != 0;
    return DFSPH_OK;
}

// prime.py:53-115 (guess: k = p / rho0 per particle, device)
int32_t dfsph_prime(void* h, float* q, float* qd, const float* guess, double dt, double tol, int32_t max_iterations,
                    int32_t* iters, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    if (!(dt > 0.0)) return fail(DFSPH_EINVAL, "dt must be > 0");
    Handle* H = as_handle(h);
    Ctx& c = H->c;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (c.n == 0) {
        if (iters) *iters = 0;
        return DFSPH_OK;
    }
    bind_state(H, q, qd, nullptr);
    HIPCHK(hipMemcpyAsync(H->qd_saved, qd, 12 * size_t(c.n), hipMemcpyDeviceToDevice, st));
    int rc;
    if ((rc = phase_neighbors(H, st))) return rc;
    if ((rc = phase_slots(H, st))) return rc;
    if ((rc = phase_density(H, st))) return rc;
    k_scale_f32_rev<<<blocks_for(c.n), 256, 0, st>>>(c.kv, guess, float(dt), c.n);
    if ((rc = phase_nonpressure(H, dt, st))) return rc;
    if ((rc = run_visc(H, dt, st, false))) return rc;  // prime.py:78: _nonpressure(predict=False)
    SolveOverride ov;
    ov.tol = tol;
    ov.max_it = max_iterations;
    ov.warm = 1.0;
    ov.min_nb = 0.0;
    ov.err_factor = 1.0e9;
    if ((rc = phase_solve(H, 0, dt, -1, st, ov))) return rc;
    HIPCHK(hipMemcpyAsync(qd, H->qd_saved, 12 * size_t(c.n), hipMemcpyDeviceToDevice, st));
    k_fill_f32<<<blocks_for(12 * int64_t(c.n)), 256, 0, st>>>(c.s_imp, 0.0f, 12 * int64_t(c.n));
    k_scale_f32<<<blocks_for(c.n), 256, 0, st>>>(c.kd, c.kv, float(dt), c.n);
    k_fill_f32<<<blocks_for(c.n), 256, 0, st>>>(c.kv, 0.0f, c.n);
    LAUNCH_CHECK();
    std::vector<int> it(c.W);
    HIPCHK(hipMemcpyAsync(it.data(), c.it_div, c.W * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (iters) *iters = *std::max_element(it.begin(), it.end());
    return DFSPH_OK;
}

int32_t dfsph_set_warm_start_model(void* h, const void* weights, void* stream) {
    (void)stream;
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    Handle* H = as_handle(h);
    H->mlp_params = static_cast<const float*>(weights);
    apply_learned(H);
    return DFSPH_OK;
}

// training-data logging of dfsph_step: features [n][16] and targets [n][2] (device; NULL stops logging)
int32_t dfsph_set_logging(void* h, float* features, float* targets) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    Handle* H = as_handle(h);
    H->log_feat = features;
    H->log_targets = targets;
    return DFSPH_OK;
}

// learned warm start as given kappa fields [n] (device, NULL = the reference rule) for the divergence and the
// density solve
int32_t dfsph_set_warm_start_kappa(void* h, const float* kappa_div, const float* kappa_den) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    Handle* H = as_handle(h);
    H->learned_div = const_cast<float*>(kappa_div);
    H->learned_den = const_cast<float*>(kappa_den);
    return DFSPH_OK;
}

int32_t dfsph_set_check_interval(void* h, int32_t every) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    if (every < 1 || every > 64) return fail(DFSPH_EINVAL, "check interval must be in [1, 64]");
    as_handle(h)->check_every = every;
    return DFSPH_OK;
}

// tuning options (none changes a result bit, except which solves the learned warm start feeds)
int32_t dfsph_set_option(void* h, int32_t option, int32_t value) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    Handle* H = as_handle(h);
    switch (option) {
        case DFSPH_OPT_CHECK_INTERVAL:
            return dfsph_set_check_interval(h, value);
        case DFSPH_OPT_GRAPHS:
            return dfsph_set_graphs(h, value);
        case DFSPH_OPT_PAIR_CACHE:
            if (value < 0 || value > DFSPH_PAIR_CACHE_ALL)
                return fail(DFSPH_EINVAL, "pair cache: a mask of DFSPH_PAIR_CACHE_* bits");
            H->c.pair_cache = value;
            return DFSPH_OK;
        case DFSPH_OPT_BUILD_BLOCK:
            if (value != 64 && value != 128 && value != 256)
                return fail(DFSPH_EINVAL, "build block: 64, 128 or 256");
            if (value == 256 && H->c.wide) return fail(DFSPH_EINVAL, "build block 256 needs uint16 rows (no world over 65,536)");
            H->build_block = value;
            return DFSPH_OK;
        case DFSPH_OPT_SPARSE_CORRECTION:
            if (value < 0 || value > 1) return fail(DFSPH_EINVAL, "sparse correction: 0 or 1");
            H->c.sparse_corr = value;
            return DFSPH_OK;
        case DFSPH_OPT_PACKED_ACCEL:
            if (value < 0 || value > 1) return fail(DFSPH_EINVAL, "packed accel: 0 or 1");
            if (value == 1 && !H->packed_ok)
                return fail(DFSPH_EINVAL, "packed accel needs one particle mass per world");
            if (H->c.packed != value) {
                // bring the layout being switched to up to date, so a phase call before the next step reads it
                if (value) k_pack_sync<<<blocks_for(H->c.n), 256>>>(H->c);
                else k_unpack_acc<<<blocks_for(H->c.n), 256>>>(H->c);
                LAUNCH_CHECK();
                HIPCHK(hipDeviceSynchronize());
                H->c.packed = value;
            }
            return DFSPH_OK;
        case DFSPH_OPT_LEARNED_SOLVES:
            if (value < 0 || value > 3) return fail(DFSPH_EINVAL, "learned solves: bit 0 divergence, bit 1 density");
            H->learned_solves = value;
            apply_learned(H);
            return DFSPH_OK;
        case DFSPH_OPT_DEVICE_LOOP:
            if (value < 0 || value > 1) return fail(DFSPH_EINVAL, "device loop: 0 or 1");
            H->device_loop = value;
            return DFSPH_OK;
        case DFSPH_OPT_BUILD_STAGED:
            if (value < 0 || value > 1) return fail(DFSPH_EINVAL, "staged build: 0 or 1");
            H->build_staged = value;
            return DFSPH_OK;
        case DFSPH_OPT_LEARNED_GATE:
            if (value < 0 || value > 1) return fail(DFSPH_EINVAL, "learned gate: 0 or 1");
            H->learned_gate = value;
            return DFSPH_OK;
        case DFSPH_OPT_HALO:
            if (value < 0 || value > 1) return fail(DFSPH_EINVAL, "halo: 0 or 1");
            if (H->c.use_halo != value) {
                H->c.use_halo = value;
                // the halos of the current rows do not exist yet: gather directly until the next list build
                int nc = std::max(H->n_chunks, 1);
                k_fill_i32<<<blocks_for(nc), 256>>>(H->c.h_count, -1, nc);
                LAUNCH_CHECK();
                HIPCHK(hipDeviceSynchronize());
            }
            return DFSPH_OK;
        default:
            return fail(DFSPH_EINVAL, "unknown option");
    }
}
int32_t dfsph_get_option(void* h, int32_t option, int32_t* value) {
    if (!h || !value) return fail(DFSPH_EINVAL, "null argument");
    Handle* H = as_handle(h);
    switch (option) {
        case DFSPH_OPT_CHECK_INTERVAL:
            *value = H->check_every;
            return DFSPH_OK;
        case DFSPH_OPT_GRAPHS:
            *value = H->use_graphs ? 1 : 0;
            return DFSPH_OK;
        case DFSPH_OPT_PAIR_CACHE:
            *value = H->c.pair_cache;
            return DFSPH_OK;
        case DFSPH_OPT_LEARNED_SOLVES:
            *value = H->learned_solves;
            return DFSPH_OK;
        case DFSPH_OPT_SPARSE_CORRECTION:
            *value = H->c.sparse_corr;
            return DFSPH_OK;
        case DFSPH_OPT_PACKED_ACCEL:
            *value = H->c.packed;
            return DFSPH_OK;
        case DFSPH_OPT_BUILD_BLOCK:
            *value = H->build_block;
            return DFSPH_OK;
        case DFSPH_OPT_HALO:
            *value = H->c.use_halo;
            return DFSPH_OK;
        case DFSPH_OPT_DEVICE_LOOP:
            *value = H->device_loop;
            return DFSPH_OK;
        case DFSPH_OPT_BUILD_STAGED:
            *value = H->build_staged;
            return DFSPH_OK;
        case DFSPH_OPT_LEARNED_GATE:
            *value = H->learned_gate;
            return DFSPH_OK;
        default:
            return fail(DFSPH_EINVAL, "unknown option");
    }
}

// per-launch timing of a 1 / sample pseudo-random subset of the launches of the kernels in `mask` (bit k: kernel
// id k; 0 stops): enabling resets the totals
int32_t dfsph_set_profiling(void* h, int32_t mask, int32_t sample) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    if (sample < 1) return fail(DFSPH_EINVAL, "profiling sample must be >= 1");
    Handle* H = as_handle(h);
    Prof& P = H->prof;
    int rc = prof_harvest(H, true);
    if (rc) return rc;
    const size_t nb = size_t(2) * size_t(std::max(H->n_chunks, 1)) * 4 * sizeof(unsigned long long);
    if (mask && !P.d_bytes) {
        void* q = nullptr;
        HIPCHK(hipMalloc(&q, nb));
        H->owned.push_back(q);
        P.d_bytes = static_cast<unsigned long long*>(q);
    }
    if (mask) {
        HIPCHK(hipMemset(P.d_bytes, 0, nb));
        for (int k = 0; k < KID_COUNT; ++k) {
            P.total_ms[k] = 0.0;
            P.count[k] = 0;
            P.launches[k] = 0;
        }
    }
    P.on = mask != 0;
    P.mask = uint32_t(mask);
    P.sample = sample;
    return DFSPH_OK;
}

// launches, summed duration [ms] and counted algorithmic bytes of kernel `kid` since profiling was enabled (syncs;
// bytes are counted by k_accel and k_apply only, 0 for the others)
int32_t dfsph_profile_read(void* h, int32_t kid, int32_t* count, double* total_ms, double* alg_bytes) {
    if (!h || !count || !total_ms) return fail(DFSPH_EINVAL, "null argument");
    if (kid < 0 || kid >= KID_COUNT) return fail(DFSPH_EINVAL, "unknown kernel id");
    Handle* H = as_handle(h);
    HIPCHK(hipDeviceSynchronize());
    int rc = prof_harvest(H, true);
    if (rc) return rc;
    *count = H->prof.count[kid];
    *total_ms = H->prof.total_ms[kid];
    if (alg_bytes) {
        *alg_bytes = 0.0;
        if (H->prof.d_bytes && (kid == KID_ACCEL || kid == KID_APPLY)) {
            size_t m = size_t(std::max(H->n_chunks, 1)) * 4;
            std::vector<unsigned long long> v(m);
            HIPCHK(hipMemcpy(v.data(), H->prof.d_bytes + size_t(kid) * m, m * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost));
            unsigned long long b = 0;
            for (unsigned long long x : v) b += x;
            *alg_bytes = double(b);
        }
    }
    return DFSPH_OK;
}

// the rows' view in the reference's layout (allocated on the first request), refreshed from the quads (to_view) or
// written back into them
static int rows_view(Handle* H, bool to_view) {
    Ctx& c = H->c;
    const int64_t count = int64_t(c.cap) * std::max(c.n, 1);
    if (!H->nbr_view) {
        if (int rc = dalloc(H, &H->nbr_view, count)) return rc;
        if (int rc = dalloc(H, &H->nbr_hi_view, c.wide ? count : 1)) return rc;
    }
    HIPCHK(hipDeviceSynchronize());
    k_rows_view<<<blocks_for(int64_t(c.cap) * c.n), 256>>>(c, H->nbr_view, H->nbr_hi_view, to_view ? 1 : 0);
    LAUNCH_CHECK();
    HIPCHK(hipDeviceSynchronize());
    return DFSPH_OK;
}

int32_t dfsph_get_buffer(void* h, int32_t which, void** ptr, int64_t* nb) {
    if (!h || !ptr || !nb) return fail(DFSPH_EINVAL, "null argument");
    Handle* H = as_handle(h);
    Ctx& c = H->c;
    int64_t n = std::max(c.n, 1), W = c.W;
#define B(p, bytes)                   \
    {                                 \
        *ptr = (void*)(p);            \
        *nb = int64_t(bytes);         \
        return DFSPH_OK;              \
    }
    switch (which) {
        case DFSPH_BUF_NBR:
            if (int rc = rows_view(H, true)) return rc;
            B(H->nbr_view, H->nbr_bytes)
        case DFSPH_BUF_NBR_HI:
            if (int rc = rows_view(H, true)) return rc;
            B(H->nbr_hi_view, H->nbr_hi_bytes)
        case DFSPH_BUF_COUNT: B(c.count, 4 * n)
        case DFSPH_BUF_GROUP: B(c.group, 4 * n)
        case DFSPH_BUF_SLOT_SHAPE: B(c.s_shape, 16 * n)
        case DFSPH_BUF_SLOT_PSI: B(c.s_psi, 16 * n)
        case DFSPH_BUF_SLOT_GPSI: B(c.s_gpsi, 48 * n)
        case DFSPH_BUF_SLOT_XB: B(c.s_xb, 48 * n)
        case DFSPH_BUF_SLOT_VB: B(c.s_vb, 48 * n)
        case DFSPH_BUF_SLOT_PHI: B(c.s_phi, 16 * n)
        case DFSPH_BUF_SLOT_XBL: B(c.s_xbl, 48 * n)
        case DFSPH_BUF_SLOT_NBL: B(c.s_nbl, 48 * n)
        case DFSPH_BUF_SLOT_IMPULSE: B(c.s_imp, 48 * n)
        case DFSPH_BUF_RATIO: B(c.ratio, 4 * n)
        case DFSPH_BUF_ALPHA: B(c.alpha, 4 * n)
        case DFSPH_BUF_NEFF: B(c.neff, 4 * n)
        case DFSPH_BUF_ACTIVE: B(c.active, 4 * n)
        case DFSPH_BUF_REG: B(c.reg, 4 * n)
        case DFSPH_BUF_SRC: B(c.src, 4 * n)
        case DFSPH_BUF_KAPPA: B(c.kappa, 4 * n)
        case DFSPH_BUF_ACC:
            if (c.packed) {  // the packed records hold the acceleration: unpack it
                k_unpack_acc<<<blocks_for(c.n), 256>>>(c);
                LAUNCH_CHECK();
                HIPCHK(hipDeviceSynchronize());
            }
            B(c.acc4, 16 * n)
        case DFSPH_BUF_ERR: B(c.err, 4 * n)
        case DFSPH_BUF_LOWER: B(c.lower, 4 * n)
        case DFSPH_BUF_KV: B(c.kv, 4 * n)
        case DFSPH_BUF_KD: B(c.kd, 4 * n)
        case DFSPH_BUF_VTMP: B(c.vtmp, 12 * n)
        case DFSPH_BUF_CONVERGED: B(c.converged, 4 * W)
        case DFSPH_BUF_ITERS: B(c.iters, 4 * W)
        case DFSPH_BUF_ERR_MEAN: B(c.err_mean, 4 * W)
        case DFSPH_BUF_ERR_MAX: B(c.err_max, 4 * W)
        case DFSPH_BUF_N_ACTIVE: B(c.n_active, 8 * W)
        case DFSPH_BUF_TOL_SCALE: B(c.tol_scale, 4 * W)
        case DFSPH_BUF_OVERFLOW: B(c.st_overflow, 4 * W)
        case DFSPH_BUF_MAX_SPEED: B(c.st_max_speed, 4 * W)
        case DFSPH_BUF_ITERS_DEN: B(c.it_den, 4 * W)
        case DFSPH_BUF_ITERS_DIV: B(c.it_div, 4 * W)
        case DFSPH_BUF_DEN_ERR: B(c.st_den_err, 4 * W)
        case DFSPH_BUF_DIV_ERR: B(c.st_div_err, 4 * W)
        case DFSPH_BUF_SHAPE_METER: B(c.meter, 24 * int64_t(std::max(c.S, 1)))
        case DFSPH_BUF_FEATURES: B(c.feat, 64 * n)
        case DFSPH_BUF_MASS_INITIAL: B(H->em.mass0, 8 * W)
        case DFSPH_BUF_MASS_IN: B(H->em.mass_in, 8 * W)
        case DFSPH_BUF_MASS_OUT: B(H->em.mass_out, 8 * W)
        case DFSPH_BUF_MASS_ACTIVE: B(H->em.mass_act, 8 * W)
        case DFSPH_BUF_ESCAPED: B(H->em.escaped, 4 * W)
        case DFSPH_BUF_EMIT_OVERFLOW: B(H->em.emit_overflow, 4 * W)
        case DFSPH_BUF_FLUX_E: B(H->em.flux_e, 8 * W)
        case DFSPH_BUF_FLUX_G: B(H->em.flux_g, 8 * W)
        case DFSPH_BUF_EMIT_T:
            if (H->em.E == 0) return fail(DFSPH_EINVAL, "no emitters");
            B(H->em.t_acc, 4 * int64_t(H->em.E))
        case DFSPH_BUF_EMIT_CURSOR:
            if (H->em.E == 0) return fail(DFSPH_EINVAL, "no emitters");
            B(H->em.cursor, 4 * int64_t(H->em.E))
        case DFSPH_BUF_EMIT_ENABLED:
            if (H->em.E == 0) return fail(DFSPH_EINVAL, "no emitters");
            B(H->em.enabled, 4 * int64_t(H->em.E))
        case DFSPH_BUF_FLAGS: B(c.flags, 4 * int64_t(c.n))
        case DFSPH_BUF_PHASE: B(c.phase, int64_t(c.n))
        case DFSPH_BUF_PW_EXCH: B(c.pw_exch, 24 * int64_t(std::max(H->n_wall, 1)))
        case DFSPH_BUF_PW_WORK: B(c.pw_work, 8 * W)
        case DFSPH_BUF_VISC_ITERS:
            if (!H->P.viscosity_implicit) return fail(DFSPH_EINVAL, "no implicit viscosity");
            B(c.v_iters, 4 * W)
        case DFSPH_BUF_CSF_SURF:
            if (!c.csf) return fail(DFSPH_EINVAL, "no csf surface tension");
            B(c.t_surf, 4 * n)
        case DFSPH_BUF_CSF_KAPPA:
            if (!c.csf) return fail(DFSPH_EINVAL, "no csf surface tension");
            B(c.t_kappa, 4 * n)
        case DFSPH_BUF_SKIN_FLAG:
            if (c.skin <= 0.0f) return fail(DFSPH_EINVAL, "no Verlet skin");
            B(c.skin_flag, 4 * W)
        case DFSPH_BUF_PW_ROW_COUNT:
            if (c.NW == 0) return fail(DFSPH_EINVAL, "no particle walls");
            B(c.wr_n, 4 * n)
        case DFSPH_BUF_HALO_COUNT: B(c.h_count, 4 * int64_t(std::max(H->n_chunks, 1)))
        default: return fail(DFSPH_EINVAL, "unknown buffer id");
    }
#undef B
}

int32_t dfsph_read_buffer(void* h, int32_t which, void* dst, int64_t nb) {
    void* p;
    int64_t have;
    int rc = dfsph_get_buffer(h, which, &p, &have);
    if (rc) return rc;
    if (nb > have) return fail(DFSPH_ECAPACITY, "read beyond the buffer");
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(dst, p, size_t(nb), hipMemcpyDeviceToHost));
    return DFSPH_OK;
}

int32_t dfsph_write_buffer(void* h, int32_t which, const void* src, int64_t nb) {
    void* p;
    int64_t have;
    int rc = dfsph_get_buffer(h, which, &p, &have);
    if (rc) return rc;
    if (nb > have) return fail(DFSPH_ECAPACITY, "write beyond the buffer");
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(p, src, size_t(nb), hipMemcpyHostToDevice));
    Ctx& c = as_handle(h)->c;
    if (which == DFSPH_BUF_ACC && c.packed) {  // into the packed records the passes read
        k_pack_sync<<<blocks_for(c.n), 256>>>(c);
        LAUNCH_CHECK();
        HIPCHK(hipDeviceSynchronize());
    }
    if (which == DFSPH_BUF_NBR || which == DFSPH_BUF_NBR_HI) return rows_view(as_handle(h), false);  // into the quads
    return DFSPH_OK;
}

int32_t dfsph_zero_stats(void* h, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    Ctx& c = as_handle(h)->c;
    hipStream_t st = static_cast<hipStream_t>(stream);
    for (int* p : {c.st_overflow, c.st_projected, c.st_swept, c.st_nonfinite}) k_fill_i32<<<blocks_for(c.W), 256, 0, st>>>(p, 0, c.W);
    k_fill_i32<<<blocks_for(6 * int64_t(std::max(c.S, 1))), 256, 0, st>>>(reinterpret_cast<int*>(c.meter), 0,
                                                                          6 * int64_t(std::max(c.S, 1)));
    LAUNCH_CHECK();
    return DFSPH_OK;
}

int32_t dfsph_report(void* h, DfsphReport* out, void* stream) {
    if (!h || !out) return fail(DFSPH_EINVAL, "null argument");
    Ctx& c = as_handle(h)->c;
    hipStream_t st = static_cast<hipStream_t>(stream);
    size_t W = size_t(c.W);
    struct Item {
        void* dst;
        const void* src;
        size_t bytes;
    } items[] = {
        {out->density_err, c.st_den_err, 4 * W},   {out->divergence_err, c.st_div_err, 4 * W},
        {out->iters_density, c.it_den, 4 * W},     {out->iters_divergence, c.it_div, 4 * W},
        {out->neighbor_overflow, c.st_overflow, 4 * W}, {out->max_speed, c.st_max_speed, 4 * W},
        {out->density_err_max, c.st_den_err_max, 4 * W}, {out->density_tol, c.st_den_tol, 4 * W},
        {out->projected, c.st_projected, 4 * W},   {out->swept, c.st_swept, 4 * W},
        {out->nonfinite, c.st_nonfinite, 4 * W},   {out->n_active, c.n_active, 8 * W},
    };
    for (auto& it : items)
        if (it.dst) HIPCHK(hipMemcpyAsync(it.dst, it.src, it.bytes, hipMemcpyDeviceToHost, st));
    int fault = 0;
    HIPCHK(hipMemcpyAsync(&fault, c.loop_fault, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (fault) return fail(DFSPH_EDEVICE, "device solve loop: a grid barrier wait ran out (code " +
                                              std::to_string(fault) + ")");
    return DFSPH_OK;
}

// emitters.py:32-255 EmitterTable / EmitterDevice and runtime.py:197-304 EmitterManager (add_emitter, add_sink,
// _upload, carry_from): the whole table, validated and copied to the device; the clock, cursor and enabled flag of
// the first min(old, new) emitters carry over.
int32_t dfsph_set_emitters(void* h, const DfsphEmitters* t) {
    if (!h || !t) return fail(DFSPH_EINVAL, "null argument");
    Handle* H = as_handle(h);
    Ctx& c = H->c;
    const int W = c.W, E = t->emitter_count, K = t->sink_count;
    if (E < 0 || K < 0) return fail(DFSPH_EINVAL, "dfsph_set_emitters: negative count");
    std::vector<int> ws(W + 1);
    HIPCHK(hipMemcpy(ws.data(), c.ws, sizeof(int) * (W + 1), hipMemcpyDeviceToHost));
    if (E > 0) {
        if (!t->world || !t->xform || !t->velocity || !t->interval || !t->offset_start || !t->pool_start)
            return fail(DFSPH_EINVAL, "dfsph_set_emitters: null emitter array");
        if (t->offset_start[0] != 0 || t->pool_start[0] != 0)
            return fail(DFSPH_EINVAL, "dfsph_set_emitters: offset_start[0] and pool_start[0] must be 0");
        for (int e = 0; e < E; ++e) {
            int w = t->world[e];
            if (w < 0 || w >= W) return fail(DFSPH_EINVAL, "add_emitter: world " + std::to_string(w) + " out of range");
            if (!(t->interval[e] > 0.0f)) return fail(DFSPH_EINVAL, "add_emitter: interval must be > 0");
            if (t->offset_start[e + 1] < t->offset_start[e] || t->pool_start[e + 1] < t->pool_start[e])
                return fail(DFSPH_EINVAL, "dfsph_set_emitters: offset_start and pool_start must be ascending");
            for (int k = t->pool_start[e]; k < t->pool_start[e + 1]; ++k) {
                int i = t->pool_ids[k];
                if (i < ws[w] || i >= ws[w + 1])
                    return fail(DFSPH_EINVAL, "dfsph_set_emitters: pool id " + std::to_string(i) +
                                                  " is not a particle of the emitter's world");
            }
        }
    }
    for (int k = 0; k < K; ++k)
        if (t->sink_world[k] < 0 || t->sink_world[k] >= W)
            return fail(DFSPH_EINVAL, "add_sink: world " + std::to_string(t->sink_world[k]) + " out of range");
    HIPCHK(hipDeviceSynchronize());
    EmitCtx& em = H->em;
    const int E_old = em.E;
    std::vector<void*> owned;
    auto up = [&](auto** dst, const auto* src, int64_t count) -> int {
        using Tv = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
        void* p = nullptr;
        size_t bytes = size_t(std::max<int64_t>(count, 1)) * sizeof(Tv);
        if (hipMalloc(&p, bytes) != hipSuccess) return fail(DFSPH_EDEVICE, "hipMalloc failed");
        owned.push_back(p);
        if (hipMemset(p, 0, bytes) != hipSuccess) return fail(DFSPH_EDEVICE, "hipMemset failed");
        if (src && count > 0 && hipMemcpy(p, src, size_t(count) * sizeof(Tv), hipMemcpyHostToDevice) != hipSuccess)
            return fail(DFSPH_EDEVICE, "hipMemcpy failed");
        *dst = static_cast<std::remove_pointer_t<decltype(dst)>>(p);
        return DFSPH_OK;
    };
    EmitCtx ne = em;  // the ledgers stay
    int rc = DFSPH_OK;
    // sinks grouped by world (the retire test is "in any sink of the world", so their order does not matter)
    std::vector<int> sstart(W + 1, 0), sorder(K);
    for (int k = 0; k < K; ++k) sstart[t->sink_world[k] + 1] += 1;
    for (int w = 0; w < W; ++w) sstart[w + 1] += sstart[w];
    {
        std::vector<int> fillp(sstart.begin(), sstart.end() - 1);
        for (int k = 0; k < K; ++k) sorder[fillp[t->sink_world[k]]++] = k;
    }
    std::vector<float> sx(7 * size_t(K)), sh(3 * size_t(K));
    for (int k = 0; k < K; ++k) {
        std::copy(t->sink_xform + 7 * sorder[k], t->sink_xform + 7 * sorder[k] + 7, sx.begin() + 7 * k);
        std::copy(t->sink_half + 3 * sorder[k], t->sink_half + 3 * sorder[k] + 3, sh.begin() + 3 * k);
    }
    const int n_off = E > 0 ? t->offset_start[E] : 0, n_pool = E > 0 ? t->pool_start[E] : 0;
    int *d_world = nullptr, *d_off_start = nullptr, *d_pool_start = nullptr, *d_pool = nullptr, *d_sstart = nullptr;
    float *d_xf = nullptr, *d_vel = nullptr, *d_int = nullptr, *d_off = nullptr, *d_sx = nullptr, *d_sh = nullptr;
    if ((rc = up(&d_world, t->world, E)) || (rc = up(&d_xf, t->xform, 7 * int64_t(E))) ||
        (rc = up(&d_vel, t->velocity, 3 * int64_t(E))) || (rc = up(&d_int, t->interval, E)) ||
        (rc = up(&d_off_start, t->offset_start, E > 0 ? E + 1 : 0)) || (rc = up(&d_off, t->offsets, 3 * int64_t(n_off))) ||
        (rc = up(&d_pool_start, t->pool_start, E > 0 ? E + 1 : 0)) || (rc = up(&d_pool, t->pool_ids, n_pool)) ||
        (rc = up(&ne.t_acc, (const float*)nullptr, E)) || (rc = up(&ne.cursor, (const int*)nullptr, E)) ||
        (rc = up(&ne.enabled, (const int*)nullptr, E)) || (rc = up(&d_sstart, sstart.data(), W + 1)) ||
        (rc = up(&d_sx, sx.data(), 7 * int64_t(K))) || (rc = up(&d_sh, sh.data(), 3 * int64_t(K)))) {
        for (void* p : owned) (void)hipFree(p);
        return rc;
    }
    std::vector<int> on(std::max(E, 1), 1);
    HIPCHK(hipMemcpy(ne.enabled, on.data(), sizeof(int) * std::max(E, 1), hipMemcpyHostToDevice));
    const int carry = std::min(E_old, E);  // EmitterDevice.carry_from
    if (carry > 0) {
        HIPCHK(hipMemcpy(ne.t_acc, em.t_acc, sizeof(float) * carry, hipMemcpyDeviceToDevice));
        HIPCHK(hipMemcpy(ne.cursor, em.cursor, sizeof(int) * carry, hipMemcpyDeviceToDevice));
        HIPCHK(hipMemcpy(ne.enabled, em.enabled, sizeof(int) * carry, hipMemcpyDeviceToDevice));
    }
    ne.E = E;
    ne.has_bounds = t->has_bounds != 0;
    ne.lo = f3{t->bounds_lo[0], t->bounds_lo[1], t->bounds_lo[2]};
    ne.hi = f3{t->bounds_hi[0], t->bounds_hi[1], t->bounds_hi[2]};
    ne.e_world = d_world;
    ne.e_xform = d_xf;
    ne.e_vel = d_vel;
    ne.e_interval = d_int;
    ne.off_start = d_off_start;
    ne.off_local = d_off;
    ne.pool_start = d_pool_start;
    ne.pool_ids = d_pool;
    ne.sink_start = d_sstart;
    ne.s_xform = d_sx;
    ne.s_half = d_sh;
    for (void* p : H->emit_owned) (void)hipFree(p);
    H->emit_owned = owned;
    H->em = ne;
    H->sink_count = K;
    H->emit_active = E > 0 || K > 0 || ne.has_bounds;
    c.speed_later = H->emit_active;
    drop_graphs(H);  // captured launches hold the previous tables
    return DFSPH_OK;
}

// SolverDFSPH.set_emitter_enabled (solver.py:438-447): a closed emitter emits nothing and its clock stops
int32_t dfsph_set_emitter_enabled(void* h, int32_t e, int32_t on) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    Handle* H = as_handle(h);
    if (e < 0 || e >= H->em.E)
        return fail(DFSPH_EINVAL, "set_emitter_enabled: emitter " + std::to_string(e) + " does not exist (" +
                                      std::to_string(H->em.E) + " emitters)");
    int v = on != 0;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(H->em.enabled + e, &v, sizeof(int), hipMemcpyHostToDevice));
    return DFSPH_OK;
}

// the step's last phase on its own: sinks, escape box, emitters and the speed statistics
int32_t dfsph_emitters_step(void* h, float* q, float* qd, double dt, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    if (!(dt > 0.0)) return fail(DFSPH_EINVAL, "dt must be > 0");
    Handle* H = as_handle(h);
    bind_state(H, q, qd, nullptr);
    return phase_emitters(H, dt, static_cast<hipStream_t>(stream));
}

// SolverDFSPH.retire (solver.py:428-431, ports.py:187-189): mask [n] int32 on the device, non-zero = retire
int32_t dfsph_retire(void* h, const int32_t* mask, float* q, float* qd, void* stream) {
    if (!h || !mask) return fail(DFSPH_EINVAL, "null argument");
    Handle* H = as_handle(h);
    bind_state(H, q, qd, nullptr);
    if (H->c.n == 0) return DFSPH_OK;
    k_retire_mask<<<blocks_for(H->c.n), 256, 0, static_cast<hipStream_t>(stream)>>>(H->c, H->em, mask);
    LAUNCH_CHECK();
    return DFSPH_OK;
}

// runtime.py:126-130 _active_mass into DFSPH_BUF_MASS_ACTIVE (the flags and phase of the model as they are now)
int32_t dfsph_active_mass(void* h, void* stream) {
    if (!h) return fail(DFSPH_EINVAL, "null handle");
    Handle* H = as_handle(h);
    Ctx& c = H->c;
    hipStream_t st = static_cast<hipStream_t>(stream);
    k_fill_i32<<<blocks_for(2 * int64_t(c.W)), 256, 0, st>>>(reinterpret_cast<int*>(H->em.mass_act), 0, 2 * int64_t(c.W));
    if (c.n > 0) k_active_mass<<<H->n_chunks, 256, 0, st>>>(c, H->em.mass_act);
    LAUNCH_CHECK();
    return DFSPH_OK;
}

}  // extern "C"
