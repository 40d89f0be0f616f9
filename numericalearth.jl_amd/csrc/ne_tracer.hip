// ne_tracer.hip — WENO5 tracer-advection tendency on the immersed LatLon grid, fused
// with the top flux boundary condition and the split Runge-Kutta stage update.
//
// Tiling (gfx950): a block of 64·TR work-items = TR wavefronts (TR = 4 for one model) owns
// a tile of 63 × TR columns and marches up the active levels. At each level the tile's T
// and S plane plus a 3-cell halo (69 × (TR+6) cells) is staged in LDS; every x- and y-face
// flux of the tile is reconstructed once (lane ℓ of wavefront r owns the west face of
// column ℓ and the south face of row r, lane 63 only a face; T, then S), written to LDS,
// and differenced by the owning cell. The vertical face flux at the top of level k is
// carried in registers as the bottom flux of level k+1, with the column's vertical stencil
// in a register window. Row-dependent values are wavefront-uniform (one wavefront per row)
// and kept in scalar registers. Tiles are ordered so that each XCD (blocks are dealt
// round-robin to the 8 XCDs) receives a contiguous band of tiles, keeping halo re-reads
// inside one L2. The level loop is software-pipelined: the inputs of level k+1 are loaded
// while level k is computed, so memory requests stay in flight through the arithmetic.
// Register pressure sets the occupancy (LDS does not): the FP64 kernel runs 4 wavefronts per
// SIMD, the FP32-storage kernel 7 (5 with the two-colour forcing); instances without the
// two-colour forcing carry none of its registers.
#include "ne_internal.h"
#include "numericalearth_hip_testing.h"

#include <algorithm>
#include <type_traits>

#include <mutex>

// Block partials of the allocation-free budget (ne_budget_f64_noalloc): one (sum,
// compensation) pair per 64×4 block of columns, NE_BUDGET_NOALLOC_MAX_BLOCKS pairs (1 MiB).
// Each device holds its own copy for the life of the process. In host code the name is a
// stand-in for the symbol, not a device address: kernels are given the address that
// hipGetSymbolAddress returns for the current device (ne::budget_noalloc_scratch).
static __device__ __align__(256) double scratch[2 * NE_BUDGET_NOALLOC_MAX_BLOCKS];

namespace ne {

namespace {

constexpr int TW = 64;            // work-items per tile row (x faces per row)
constexpr int TC = 63;            // cells per tile row
#ifndef NE_PLAIN_TR
#define NE_PLAIN_TR 4
#endif
// Minimum wavefronts per SIMD requested for the stage kernels. The software pipeline holds
// the next level's inputs in registers, and a spilled register defeats it (a reload from
// scratch waits for every load in flight), so the targets are the largest occupancies whose
// register budget fits without spills, measured on 1440×560×10: FP64 4 (≈116–125 VGPRs;
// 5 spills and runs 30 % slower), FP32 7 (≤ 72 VGPRs, a 12-byte spill, faster than 5 or 6)
// and 5 for the FP32 instances with the two-colour forcing (≈ 93 VGPRs). The FP32 state's last
// stage, which forms its update in FP64, takes 6 (≤ 80 VGPRs).
#ifndef NE_STAGE_WAVES_PER_EU_F32
#define NE_STAGE_WAVES_PER_EU_F32 7
#endif
#ifndef NE_STAGE_WAVES_PER_EU_F32_LAST
#define NE_STAGE_WAVES_PER_EU_F32_LAST 6
#endif
#ifndef NE_STAGE_WAVES_PER_EU_F32_TWO_COLOR
#define NE_STAGE_WAVES_PER_EU_F32_TWO_COLOR 5
#endif
#ifndef NE_STAGE_WAVES_PER_EU_F64
#define NE_STAGE_WAVES_PER_EU_F64 4
#endif
template <typename Real, bool TWO_COLOR, bool LAST>
constexpr int stage_waves() {
    return sizeof(Real) == 8 ? NE_STAGE_WAVES_PER_EU_F64
           : TWO_COLOR       ? NE_STAGE_WAVES_PER_EU_F32_TWO_COLOR
           : LAST            ? NE_STAGE_WAVES_PER_EU_F32_LAST
                             : NE_STAGE_WAVES_PER_EU_F32;
}
constexpr int PLAIN_TR = NE_PLAIN_TR;  // cell rows per tile of the single-model kernel
constexpr int NXCD = 8;
constexpr int KMIN_OUTSIDE = 1 << 30;  // kmin of a column outside a bounded domain

// LDS of one tile of TR rows: the T/S plane with a 3-cell halo, the x- and y-face fluxes
// and the kmin of every plane column.
template <typename R, int TR>
struct TileShared {
    static constexpr int PXW = TC + 6;   // plane columns i0-3 … i0+65
    static constexpr int PYH = TR + 6;   // plane rows    j0-3 … j0+TR+2
    static constexpr int PN = PXW * PYH;
    R cT[PN], cS[PN];
    R fxT[TR * TW], fxS[TR * TW];
    R fyT[(TR + 1) * TC], fyS[(TR + 1) * TC];
    int km[PN];
    int kstart;
    double resid[2 * TR];   // the rows' rounding residuals of T and S (FP32 state's last stage)
};

__host__ __device__ inline int tiles_x(int Nx) { return (Nx + TC - 1) / TC; }
__host__ __device__ inline int tiles_y(int Ny, int tr) { return (Ny + tr - 1) / tr; }
inline int grid_blocks(int Nx, int Ny) {
    int n = tiles_x(Nx) * tiles_y(Ny, PLAIN_TR);
    return ((n + NXCD - 1) / NXCD) * NXCD;
}

// Face order at level k from the three thresholds of its stencil: the face is
// reconstructed with order 5 (3) if k ≥ K3 (K2), upwind if k ≥ K1, and carries no flux
// otherwise (K1 ≤ K2 ≤ K3 are the largest kmin over the 2, 4 and 6 cells around it). The
// three are packed in 10-bit fields of one register (grids have at most NE_MAX_LEVELS =
// 1023 levels, so 1023 — "never" — exceeds every level).
constexpr uint32_t NEVER_LEVEL = 1023u;
__device__ __forceinline__ int order_at(int k, uint32_t P) {
    return (k >= (int)(P & 1023u) ? 1 : 0) + (k >= (int)((P >> 10) & 1023u) ? 1 : 0) + (k >= (int)(P >> 20) ? 1 : 0);
}

__device__ __forceinline__ uint32_t thresholds(const int* km, int stride) {
    const int K1 = max(km[2 * stride], km[3 * stride]);
    const int K2 = max(K1, max(km[stride], km[4 * stride]));
    const int K3 = max(K2, max(km[0], km[5 * stride]));
    auto field = [](int K) { return K < (int)NEVER_LEVEL ? (uint32_t)K : NEVER_LEVEL; };
    return field(K1) | field(K2) << 10 | field(K3) << 20;
}
constexpr uint32_t NEVER_FACE = NEVER_LEVEL | NEVER_LEVEL << 10 | NEVER_LEVEL << 20;

// Integer hash (Wellons' "lowbias32" finaliser) for the stochastic rounding below.
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Rounding of the new FP64 state value x to the storage type. FP32 storage of the model
// state (the mixed form) rounds stochastically: to one of the two floats around x with
// probabilities that make the expected stored value x, drawn from a hash of the cell, the
// tracer and the step (`seed`), so results are reproducible and identical on every path
// that takes the same step. Round-to-nearest would drop the same fraction of an ulp every
// step wherever the increment is steady, a bias that grows linearly with the step count.
// The draw adds 29 random bits h (uniform on [0, 2²⁹)) to the 29 bits of x's significand
// that a float does not hold and clears them: |x| rounds up to the next float in magnitude
// exactly when h ≥ 2²⁹ − f, f the cleared bits, i.e. with probability f·2⁻²⁹, the distance
// to the float below over the gap (a carry into the exponent is the step to the next binade),
// and the result converts to float exactly (for |x| within the normal floats).
template <typename R>
__device__ __forceinline__ R round_state(double x, uint32_t key, uint32_t seed, bool stochastic) {
    if (sizeof(R) == 4 && stochastic) {
        constexpr uint64_t LOW = (1ull << 29) - 1;
        const uint64_t h = (uint64_t)(hash32(hash32(key) ^ seed) >> 3);
        const uint64_t b = ((uint64_t)__double_as_longlong(x) + h) & ~LOW;
        return (R)__longlong_as_double((long long)b);
    }
    return (R)x;
}

// A wave-uniform value as a scalar (the tile kernels' blocks are 64 wide, so every value that
// depends on the row only is uniform over a wavefront): keeps it out of vector registers.
__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double wave_uniform(double x) {
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

// 1/x for a positive normal x: the FP64 reciprocal refined by Newton steps (within 2 ulp), the
// hardware reciprocal in FP32 (1 ulp).
__device__ __forceinline__ double reciprocal(double x) { return quotient(1.0, x); }
__device__ __forceinline__ float reciprocal(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename T>
__device__ __forceinline__ T ldb(const T* base, uint32_t off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <typename T>
__device__ __forceinline__ void stb(T* base, uint32_t off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v;
}

// Double-double accumulation (sum, compensation): the budgets and the rounding residual.
struct DD {
    double s, c;
};

__device__ __forceinline__ DD dd_add(DD a, DD b) {
    double t = a.s + b.s;
    double e;
    if (fabs(a.s) >= fabs(b.s)) e = (a.s - t) + b.s;
    else e = (b.s - t) + a.s;
    DD r;
    r.s = t;
    r.c = (a.c + b.c) + e;
    return r;
}

// One tile (columns i0 … i0+62, rows j0 … j0+TR-1) marched over its active levels.
// MODE 0: c_out = c0 + beta_dt·G(c_in) on active cells. MODE 1: write G (0 on inactive
// cells). Every work-item of the block calls this with the same tile. Every address is a
// 32-bit byte offset from a uniform base (the host checks that each view spans < 4 GiB),
// computed once per tile; a level adds its uniform k-stride.
// BEER (with TWO_COLOR): the Beer's-law factors come from F.beer instead of being computed.
template <typename Real, int MODE, int TR, bool TWO_COLOR, bool WATCH = false, bool BEER = false>
__device__ __forceinline__ void tile_body(TileShared<Real, TR>& sh, const GridDev& g, const TracerLayout& L,
                                          const Real* __restrict__ u, const Real* __restrict__ v,
                                          const Real* __restrict__ w, const Real* T0, const Real* S0,
                                          const Real* __restrict__ Tin, const Real* __restrict__ Sin, Real* Tout,
                                          Real* Sout, const double* __restrict__ JT, const double* __restrict__ JS,
                                          const double* __restrict__ Jw, const TracerForcing& F, double beta_dt,
                                          int i0, int j0, bool stochastic = false, uint32_t seed = 0u) {
    typedef Real R;
    constexpr int PXW = TileShared<R, TR>::PXW;
    constexpr int PN = TileShared<R, TR>::PN;
    constexpr int NTHR = TW * TR;
    constexpr int NPL = (PN + NTHR - 1) / NTHR;   // plane cells staged per work-item
    constexpr uint32_t ES = sizeof(R);
    constexpr int NEVER = 1 << 30;
    const int lane = threadIdx.x;
    const int row = wave_uniform((int)threadIdx.y);   // one wavefront per row
    const int tid = lane + row * TW;
    const int Nz = g.Nz;
    const bool adv = g.advection != 0;   // single-column grids carry no advection

    // kmin of the plane columns (outside a bounded domain: never active)
    for (int n = tid; n < PN; n += NTHR) {
        int pi = n % PXW, pj = n / PXW;
        int64_t c = g.column(i0 - 3 + pi, j0 - 3 + pj);
        sh.km[n] = c >= 0 ? g.kmin[c] : KMIN_OUTSIDE;
    }
    if (tid == 0) sh.kstart = Nz;
    __syncthreads();
    const int ci = i0 + lane, cj = j0 + row;
    const bool cell = lane < TC && ci < g.Nx && cj < g.Ny;
    const int mykmin = cell ? sh.km[(row + 3) * PXW + lane + 3] : Nz;
    if (cell) atomicMin(&sh.kstart, mykmin);

    // Per-tile addressing and face thresholds (independent of the level)
    int pkm[NPL];
    uint32_t poff[NPL];
#pragma unroll
    for (int q = 0; q < NPL; q++) {
        const int n = tid + q * NTHR;
        pkm[q] = NEVER;
        poff[q] = 0;
        if (n < PN) {
            pkm[q] = sh.km[n];
            if (pkm[q] < KMIN_OUTSIDE) {
                const int gi = g.wx(i0 - 3 + n % PXW), gj = g.wy(j0 - 3 + n / PXW);
                poff[q] = (uint32_t)gi * ES + (uint32_t)gj * L.c_sy;
            }
        }
    }
    const bool xrow = cj < g.Ny;                 // this wavefront's row is inside the domain
    const bool ycol = lane < TC && ci < g.Nx;
    // x face owned by this lane (west face of column i0+lane), south y face of row j0+row,
    // north y face of the tile (last wavefront)
    const uint32_t Xk = (xrow && adv) ? thresholds(&sh.km[(row + 3) * PXW + lane], 1) : NEVER_FACE;
    const uint32_t Sk = (ycol && adv) ? thresholds(&sh.km[row * PXW + lane + 3], PXW) : NEVER_FACE;
    // The tile's north face row is shared out: wavefront 0 reconstructs it for T, wavefront 1
    // for S (a single-row tile's wavefront does both), so no wavefront carries more than one
    // extra reconstruction into the level's second barrier.
    const bool north = TR >= 2 ? row < 2 : true;
    const uint32_t Nk = (ycol && adv && north) ? thresholds(&sh.km[TR * PXW + lane + 3], PXW) : NEVER_FACE;
    const int fiw = g.wx(i0 + lane);
    const int fjs = g.wy(j0 + row);
    const int fjn = g.wy(j0 + TR);
    const int cy = ycol ? ci : 0;
    const double Axf = wave_uniform(g.dy_xface[xrow ? cj : 0]);
    const double Ays =
        wave_uniform((j0 + row <= g.Ny) ? g.dx_yface[fjs < 0 ? 0 : (fjs > g.Ny ? g.Ny : fjs)] : 0.0);
    const double Ayn =
        wave_uniform((j0 + TR <= g.Ny) ? g.dx_yface[fjn < 0 ? 0 : (fjn > g.Ny ? g.Ny : fjn)] : 0.0);
    const double Azc = wave_uniform(xrow ? g.Az[cj] : 0.0);   // used on cells only
    const uint32_t uoff = xrow ? (uint32_t)fiw * ES + (uint32_t)cj * L.u_sy : 0u;
    const uint32_t voff_s = ycol ? (uint32_t)cy * ES + (uint32_t)fjs * L.v_sy : 0u;
    const uint32_t dvn = (uint32_t)(fjn - fjs) * L.v_sy;           // north face: voff_s + dvn (uniform)
    const uint32_t coff = cell ? (uint32_t)ci * ES + (uint32_t)cj * L.c_sy : 0u;
    const uint32_t dw = (uint32_t)cj * (L.w_sy - L.c_sy);           // w of the cell: coff + dw (uniform)
    auto joff = [&]() { return (uint32_t)ci * 8u + (uint32_t)cj * L.j_sy; };   // 2-D fluxes (cells only)
    __syncthreads();
    const int kstart = sh.kstart;
    if (MODE == 1 && cell)
        for (int k = 0; k < mykmin && k < Nz; k++) {
            stb(Tout, coff + (uint32_t)k * L.c_sz, (R)0);
            stb(Sout, coff + (uint32_t)k * L.c_sz, (R)0);
        }
    // freeze watch (the step's last stage; instances without it carry none of its code)
    const bool watch = WATCH && MODE == 0 && F.freeze_flags != nullptr;   // uniform
    bool frz = false;
    // Conservative last stage of an FP32 state: the face fluxes (FP32) are differenced and the
    // update formed in FP64, so Σ V·(x − c⁰) telescopes to the boundary fluxes to FP64
    // round-off, and Σ V·(x − stored) — exact per cell, as x and the stored float are within
    // an ulp — is summed for the model's rounding residual.
    constexpr bool CONS = WATCH && MODE == 0 && sizeof(R) == 4;
    const bool resid = CONS && F.residual != nullptr;   // uniform
    double resT = 0.0, resS = 0.0;
    if (kstart >= Nz) {  // uniform: the whole tile is land
        if (watch && cell) F.freeze_flags[(int64_t)ci + (int64_t)cj * g.Nx] = 0;
        return;
    }

    // Vertical register window of the owned column: levels k-2 … k+3 while level k is
    // computed; level k+3 arrives with the level's other inputs (below).
    R zT[6], zS[6];
#pragma unroll
    for (int n = 0; n < 5; n++) {
        int kk = kstart - 2 + n;
        bool a = cell && kk >= mykmin && kk < Nz;
        zT[n] = a ? ldb(Tin, coff + (uint32_t)kk * L.c_sz) : (R)0;
        zS[n] = a ? ldb(Sin, coff + (uint32_t)kk * L.c_sz) : (R)0;
    }
    zT[5] = (R)0;
    zS[5] = (R)0;
    R FzT = (R)0, FzS = (R)0;  // flux through the bottom face of the current level

    // TwoColorRadiation: J₀ of the column and the surface height η of the Beer's-law factors
    const bool tc = TWO_COLOR && F.two_color != 0;   // instances without it carry none of its registers
    const double J0 = (tc && cell) ? ldb(F.J0, joff()) : 0.0;
    const double eta = tc ? g.zf[Nz] : 0.0;

    // Inputs of one level, loaded one level ahead (a software pipeline): each input of level
    // k+1 is loaded right after its last use at level k — the plane (staged in LDS at the start
    // of the level) once it is in LDS, the x-face velocity after the x faces, the y-face ones
    // after the y faces, w after the z face, c⁰ after the update — so every wavefront keeps
    // memory requests in flight through its arithmetic while holding one copy of each input
    // (the plane and the window's top are the only values carried twice). The loads carry no
    // condition: a value a work-item does not use (inactive or outside cells, faces of order 0,
    // levels past the top) is read from an address that is always inside the view and never
    // enters arithmetic, because face orders only admit active cells. Conditional loads would
    // make the compiler wait for every load in flight wherever their paths merge.
    const uint32_t uoff_l = Xk != NEVER_FACE ? uoff : 0u;
    const uint32_t voff_l = Sk != NEVER_FACE ? voff_s : 0u;
    const uint32_t vnoff_l = Nk != NEVER_FACE ? voff_s + dvn : 0u;
    const uint32_t woff_l = cell ? coff + dw : 0u;
    R nT[NPL], nS[NPL];
    R nu, nv, nvn, nw, nT0 = (R)0, nS0 = (R)0, nzT, nzS;
    double ndz, nzb = 0.0, nzt = 0.0;   // Δz of the level, heights of its faces (two-colour)
    auto load_plane = [&](int k) {
        const uint32_t kc = (uint32_t)k * L.c_sz;
#pragma unroll
        for (int q = 0; q < NPL; q++) {
            nT[q] = ldb(Tin, poff[q] + kc);
            nS[q] = ldb(Sin, poff[q] + kc);
        }
        const uint32_t kz = (uint32_t)min(k + 3, Nz - 1) * L.c_sz;   // the window's top at level k
        nzT = ldb(Tin, coff + kz);
        nzS = ldb(Sin, coff + kz);
        // the grid's per-level values travel with the level's other loads: read at their point of
        // use, each would wait for every load in flight (memory counters complete in order)
        ndz = g.dz[k];
        if (tc && !BEER) {
            nzb = g.zf[k];
            nzt = g.zf[k + 1];
        }
    };
    auto load_u = [&](int k) { nu = ldb(u, uoff_l + (uint32_t)k * L.u_sz); };
    auto load_v = [&](int k) {
        nv = ldb(v, voff_l + (uint32_t)k * L.v_sz);
        nvn = ldb(v, vnoff_l + (uint32_t)k * L.v_sz);
    };
    auto load_w = [&](int k) { nw = ldb(w, woff_l + (uint32_t)(k + 1) * L.w_sz); };
    auto load_c0 = [&](int k) {
        if (MODE == 0) {
            nT0 = ldb(T0, coff + (uint32_t)k * L.c_sz);
            nS0 = ldb(S0, coff + (uint32_t)k * L.c_sz);
        }
    };
    load_plane(kstart);
    load_u(kstart);
    load_v(kstart);
    load_w(kstart);
    load_c0(kstart);
    for (int k = kstart; k < Nz; k++) {
        const int kn = min(k + 1, Nz - 1);   // the level loaded ahead (the last one again at the top)
        const uint32_t kc = (uint32_t)k * L.c_sz;
        // 1. stage the plane of level k; take its velocities, c⁰ and the window's top
#pragma unroll
        for (int q = 0; q < NPL; q++) {
            const int n = tid + q * NTHR;
            if (n < PN) {
                sh.cT[n] = nT[q];
                sh.cS[n] = nS[q];
            }
        }
        const double dzd = wave_uniform(ndz), zfb = wave_uniform(nzb), zft = wave_uniform(nzt);
        zT[5] = nzT;
        zS[5] = nzS;
        __syncthreads();
        load_plane(kn);
        const R dzk = (R)dzd;

        // 2a. x face owned by this lane: cells i0+lane-3 … i0+lane+2 of row j0+row; T, then S.
        // FP64 interleaves the two tracers' reconstructions (independent chains to issue from:
        // 8 µs of 158 per stage, measured); FP32 takes one at a time, whose temporaries alone fit
        // its 7-wavefront register budget.
        constexpr int FACE_UNROLL = sizeof(R) == 8 ? 2 : 1;
        {
            const int order = order_at(k, Xk);
            const R A = (R)Axf * dzk;
            const int base = (row + 3) * PXW + lane;
#pragma unroll FACE_UNROLL
            for (int tr = 0; tr < 2; tr++) {
                const R* t = (tr ? sh.cS : sh.cT) + base;
                R f = (R)0;
                if (order > 0) {
                    f = A * nu * face_value(t[0], t[1], t[2], t[3], t[4], t[5], order, nu);
                }
                (tr ? sh.fxS : sh.fxT)[row * TW + lane] = f;
            }
        }
        load_u(kn);
        // 2b. y faces: the south face of row j0+row, and for the last wavefront the north face of the tile
        if (ycol) {
#pragma unroll FACE_UNROLL
            for (int pass = 0; pass < (TR >= 2 ? (north ? 3 : 2) : 4); pass++) {
                const int side = pass >> 1;               // 0: south face of the row, 1: north face of the tile
                const int tr = (TR >= 2 && side) ? row : (pass & 1);
                const int r0 = side ? TR : row;           // plane row of the first stencil cell
                const double Ay = side ? Ayn : Ays;
                const int order = order_at(k, side ? Nk : Sk);
                R f = (R)0;
                if (order > 0) {
                    const R vel = side ? nvn : nv;
                    const R A = (R)Ay * dzk;
                    const R* t = (tr ? sh.cS : sh.cT) + r0 * PXW + lane + 3;
                    f = A * vel * face_value(t[0], t[PXW], t[2 * PXW], t[3 * PXW], t[4 * PXW], t[5 * PXW], order, vel);
                }
                (tr ? sh.fyS : sh.fyT)[r0 * TC + lane] = f;
            }
        }
        load_v(kn);
        // 2c. z face at the top of level k (cells k-2 … k+3 of the owned column)
        const int zorder = (cell && adv) ? (k >= mykmin && k <= Nz - 2 ? 1 : 0) +
                                               (k >= mykmin + 1 && k <= Nz - 3 ? 1 : 0) +
                                               (k >= mykmin + 2 && k <= Nz - 4 ? 1 : 0)
                                         : 0;
        R FztT = (R)0, FztS = (R)0;
        if (zorder > 0) {
            FztT = (R)Azc * nw * face_value(zT[0], zT[1], zT[2], zT[3], zT[4], zT[5], zorder, nw);
            FztS = (R)Azc * nw * face_value(zS[0], zS[1], zS[2], zS[3], zS[4], zS[5], zorder, nw);
        }
        load_w(kn);
        __syncthreads();

        // 3. difference the face fluxes, add the top boundary flux, update
        if (cell && k >= mykmin) {
            typedef typename std::conditional<CONS, double, R>::type D;   // arithmetic of the update
            const int xo = row * TW + lane, yo = row * TC + lane;
            D div_T = ((D)sh.fxT[xo + 1] - (D)sh.fxT[xo]) + ((D)sh.fyT[yo + TC] - (D)sh.fyT[yo]);
            div_T = div_T + ((D)FztT - (D)FzT);
            D div_S = ((D)sh.fxS[xo + 1] - (D)sh.fxS[xo]) + ((D)sh.fyS[yo + TC] - (D)sh.fyS[yo]);
            div_S = div_S + ((D)FztS - (D)FzS);
            const D V = (D)(Azc * dzd);
            const D rV = reciprocal(V);
            D GT = -(rV * div_T);
            D GS = -(rV * div_S);
            if (tc) {
                // ε₁ ∂z J₁ + (1 − ε₁) ∂z J₂; the face below the lowest active cell carries no flux
                // (the level's two faces; uniform over the tile, recomputed rather than carried)
                double E1t, E2t, E1b, E2b;
                if constexpr (BEER) {
                    // the level's two faces from the table, at a uniform index: scalar loads, which
                    // wait on their own counter and leave the vector loads in flight alone
                    const int ku = wave_uniform(k);
                    E1b = F.beer[2 * ku];
                    E2b = F.beer[2 * ku + 1];
                    E1t = F.beer[2 * ku + 2];
                    E2t = F.beer[2 * ku + 3];
                } else {
                    E1t = exp(F.kappa1 * (zft - eta));
                    E2t = exp(F.kappa2 * (zft - eta));
                    E1b = exp(F.kappa1 * (zfb - eta));
                    E2b = exp(F.kappa2 * (zfb - eta));
                }
                const bool below = k - 1 >= mykmin;
                const double a1 = J0 * E1t, b1 = below ? J0 * E1b : 0.0;
                const double a2 = J0 * E2t, b2 = below ? J0 * E2b : 0.0;
                const double dz = dzd;
                const double dJ1dz = (a1 - b1) / dz;
                const double dJ2dz = (a2 - b2) / dz;
                GT = GT + (D)(F.first_fraction * dJ1dz + (1.0 - F.first_fraction) * dJ2dz);
            }
            if (k == Nz - 1) {
                // MultipleFluxes(Jᵀ, FreshwaterExchange{:T}) = Jᵀ + 0 on a static grid;
                // MultipleFluxes(Jˢ, FreshwaterExchange{:S}) = Jˢ + (Sᴺ Jʷ − 0)
                // (multiple_surface_fluxes.jl:27-31; ocean_simulation.jl:80-91)
                const uint32_t jo = joff();
                double FT = ldb(JT, jo) + (0.0 + 0.0);
                double FS = ldb(JS, jo) + (((double)zS[2] * ldb(Jw, jo) - 0.0) + 0.0);
                GT = GT - (D)(FT * Azc) * rV;
                GS = GS - (D)(FS * Azc) * rV;
            }
            if (MODE == 0) {
                // The update is formed in FP64 and rounded once to the storage type (for FP32
                // storage this keeps the per-stage rounding of the state at half an ulp).
                const uint32_t key = 2u * ((uint32_t)ci + (uint32_t)g.Nx * ((uint32_t)cj + (uint32_t)g.Ny * k));
                const double xT = (double)nT0 + beta_dt * (double)GT;
                const double xS = (double)nS0 + beta_dt * (double)GS;
                const R Tn = round_state<R>(xT, key, seed, stochastic);
                const R Sn = round_state<R>(xS, key + 1u, seed, stochastic);
                stb(Tout, coff + kc, Tn);
                stb(Sout, coff + kc, Sn);
                if (resid) {
                    resT += (double)V * (xT - (double)Tn);
                    resS += (double)V * (xS - (double)Sn);
                }
                // the freeze clamp's test on the stored values (freeze_column, ne_surface.hip)
                if (watch) frz = frz || (double)Tn < F.melt0 - F.liquidus_slope * (double)Sn;
            } else {
                stb(Tout, coff + kc, (R)GT);
                stb(Sout, coff + kc, (R)GS);
            }
        }
        load_c0(kn);
        FzT = FztT;
        FzS = FztS;
        // 4. slide the vertical window up one level (its new top arrives with level k+1)
#pragma unroll
        for (int n = 0; n < 5; n++) { zT[n] = zT[n + 1]; zS[n] = zS[n + 1]; }
    }
    if (watch && cell) F.freeze_flags[(int64_t)ci + (int64_t)cj * g.Nx] = frz ? 1 : 0;
    if (resid) {   // the tile's residuals: each row's by a fixed butterfly, then the rows in order
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            resT += __shfl_xor(resT, off);
            resS += __shfl_xor(resS, off);
        }
        if (lane == 0) {
            sh.resid[2 * row] = resT;
            sh.resid[2 * row + 1] = resS;
        }
        __syncthreads();
        if (tid == 0) {   // into the tile's running double-double totals
            double sT = 0.0, sS = 0.0;
            for (int r = 0; r < TR; r++) {
                sT += sh.resid[2 * r];
                sS += sh.resid[2 * r + 1];
            }
            const DD T = dd_add(DD{F.residual[0], F.residual[1]}, DD{sT, 0.0});
            const DD S = dd_add(DD{F.residual[2], F.residual[3]}, DD{sS, 0.0});
            F.residual[0] = T.s;
            F.residual[1] = T.c;
            F.residual[2] = S.s;
            F.residual[3] = S.c;
        }
    }
}

// Single model: one block per tile, tiles ordered so that each XCD (blocks are dealt
// round-robin to the 8 XCDs) receives a contiguous band of tiles.
template <typename Real, int MODE, bool TWO_COLOR, bool WATCH = false, bool BEER = false>
__global__ void __launch_bounds__(TW * PLAIN_TR)
__attribute__((amdgpu_waves_per_eu(stage_waves<Real, TWO_COLOR, WATCH && MODE == 0>())))
tracer_tile_kernel(GridDev g, TracerLayout L, const Real* u, const Real* v, const Real* w, const Real* T0,
                   const Real* S0, const Real* Tin, const Real* Sin, Real* Tout, Real* Sout, const double* JT,
                   const double* JS, const double* Jw, TracerForcing F, double beta_dt, int stochastic,
                   uint32_t seed) {
    __shared__ TileShared<Real, PLAIN_TR> sh;
    const int ntx = tiles_x(g.Nx), nty = tiles_y(g.Ny, PLAIN_TR);
    const int ntiles = ntx * nty;
    const int per = (ntiles + NXCD - 1) / NXCD;
    const int tile = (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
    if (tile >= ntiles) return;
    if (F.residual) F.residual += RESIDUAL_SLOT * tile;   // this tile's totals
    tile_body<Real, MODE, PLAIN_TR, TWO_COLOR, WATCH, BEER>(sh, g, L, u, v, w, T0, S0, Tin, Sin, Tout, Sout, JT, JS, Jw, F,
                                                     beta_dt, (tile % ntx) * TC, (tile / ntx) * PLAIN_TR,
                                                     stochastic != 0, seed);
}

// Ensemble: a persistent grid walks the (member, tile) work list. The list is cut into 8
// contiguous chunks, chunk x going to the blocks of XCD x, which stride through it; each
// member's tiles are contiguous in the list, so a member's halo re-reads stay in one L2.
// stage 1: T¹ = T + β₁Δt G(T); stage 2: T² = T + β₂Δt G(T¹); stage 3: T = T + β₃Δt G(T²).
template <typename Real, int TR, bool TWO_COLOR, bool WATCH>
__global__ void __launch_bounds__(TW * TR)
__attribute__((amdgpu_waves_per_eu(stage_waves<Real, TWO_COLOR, WATCH>())))
batch_stage_kernel(const BatchRec<Real>* __restrict__ recs, const WorkItem* __restrict__ work, int nwork,
                   int stage, double beta_dt) {
    __shared__ TileShared<Real, TR> sh;
    const int Q = (int)gridDim.x / NXCD;
    const int x = (int)blockIdx.x % NXCD, q = (int)blockIdx.x / NXCD;
    const int chunk = (nwork + NXCD - 1) / NXCD;
    const int end = min((x + 1) * chunk, nwork);
    for (int wi = x * chunk + q; wi < end; wi += Q) {
        const WorkItem it = work[wi];
        const BatchRec<Real>& rc = recs[it.member];
        const GridDev g = rc.g;
        const TracerLayout L = tracer_layout(rc);
        const int ntx = tiles_x(g.Nx);
        const Real* Tin = stage == 1 ? rc.T.p : stage == 2 ? rc.T1.p : rc.T2.p;
        const Real* Sin = stage == 1 ? rc.S.p : stage == 2 ? rc.S1.p : rc.S2.p;
        Real* Tout = stage == 1 ? rc.T1.p : stage == 2 ? rc.T2.p : rc.T.p;
        Real* Sout = stage == 1 ? rc.S1.p : stage == 2 ? rc.S2.p : rc.S.p;
        TracerForcing F = {0, 0.0, 0.0, 0.0, nullptr, nullptr, 0.0, 0.0, nullptr, nullptr};
        if (rc.p.radiative_forcing == NE_RADIATIVE_FORCING_TWO_COLOR) {
            F.two_color = 1;
            F.first_fraction = rc.p.first_color_fraction;
            F.kappa1 = rc.p.first_absorption_coefficient;
            F.kappa2 = rc.p.second_absorption_coefficient;
            F.J0 = rc.n.J_sw.p;
            F.beer = rc.beer;
        }
        if (WATCH) {   // the step's last stage flags the columns the freeze clamp visits
            F.freeze_flags = rc.freeze_flags;
            F.melt0 = rc.p.freshwater_melting_temperature;
            F.liquidus_slope = rc.p.liquidus_slope;
            if (sizeof(Real) == 4 && rc.residual) F.residual = rc.residual + RESIDUAL_SLOT * it.index;
        }
        __syncthreads();   // the previous item's reads of the LDS are complete
        tile_body<Real, 0, TR, TWO_COLOR, WATCH, TWO_COLOR>(sh, g, L, rc.u.p, rc.v.p, rc.w.p, rc.T.p, rc.S.p, Tin, Sin, Tout, Sout, rc.n.J_T.p,
                               rc.n.J_S.p, rc.n.J_w.p, F, beta_dt, (it.index % ntx) * TC, (it.index / ntx) * TR,
                               stage == 3, step_seed(rc.time));
    }
}

// ---------------------------------------------------------------------------
// Vertical velocity from continuity: one work-item per column marching up the levels,
// i fastest so that every level's u, v and w rows are read and written coalesced. The
// running w is kept in FP64 and rounded once per face to the storage type.
// ---------------------------------------------------------------------------

template <typename Real>
__global__ void __launch_bounds__(256) w_continuity_kernel(GridDev g, View3<const Real> u, View3<const Real> v,
                                                           View3<Real> w) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    if (i >= g.Nx || j >= g.Ny) return;
    const int ip = g.tx == NE_PERIODIC ? wrap_near(i + 1, g.Nx) : i + 1;
    const int jp = g.ty == NE_PERIODIC ? wrap_near(j + 1, g.Ny) : j + 1;
    const double dyx = g.dy_xface[j], dxs = g.dx_yface[j], dxn = g.dx_yface[jp], Az = g.Az[j];
    double wk = 0.0;
    w(i, j, 0) = (Real)0;
    for (int k = 0; k < g.Nz; k++) {
        const double dz = g.dz[k];
        const double Ax = dyx * dz, Ays = dxs * dz, Ayn = dxn * dz;
        const double div = (Ax * (double)u(ip, j, k) - Ax * (double)u(i, j, k)) +
                           (Ayn * (double)v(i, jp, k) - Ays * (double)v(i, j, k));
        wk = g.active(i, j, k) ? wk - div / Az : 0.0;
        w(i, j, k + 1) = (Real)wk;
    }
}

// ---------------------------------------------------------------------------
// Compensated budget: Neumaier per column, then a fixed-order double-double tree.
// ---------------------------------------------------------------------------

constexpr int BX = 64;
constexpr int BY = 4;

inline dim3 grid2d(int Nx, int Ny) { return dim3((Nx + BX - 1) / BX, (Ny + BY - 1) / BY); }

template <typename Real>
__global__ void budget_partial_kernel(GridDev g, View3<const Real> c, double weight, double* partial) {
    __shared__ double ss[BX * BY], sc[BX * BY];
    int i = blockIdx.x * BX + threadIdx.x;
    int j = blockIdx.y * BY + threadIdx.y;
    int t = threadIdx.x + threadIdx.y * BX;
    DD acc = {0.0, 0.0};
    if (i < g.Nx && j < g.Ny) {
        int k0 = g.kmin[i + (int64_t)j * g.Nx];
        for (int k = k0; k < g.Nz; k++) {
            double x = weight * (g.Az[j] * g.dz[k]) * (double)c(i, j, k);
            DD xd = {x, 0.0};
            acc = dd_add(acc, xd);
        }
    }
    ss[t] = acc.s;
    sc[t] = acc.c;
    __syncthreads();
    for (int w = BX * BY / 2; w > 0; w >>= 1) {
        if (t < w) {
            DD a = {ss[t], sc[t]}, b = {ss[t + w], sc[t + w]};
            DD r = dd_add(a, b);
            ss[t] = r.s;
            sc[t] = r.c;
        }
        __syncthreads();
    }
    if (t == 0) {
        int64_t b = blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
        partial[2 * b] = ss[0];
        partial[2 * b + 1] = sc[0];
    }
}

__global__ void budget_final_kernel(const double* partial, int64_t n, double* out) {
    __shared__ double ss[256], sc[256];
    int t = threadIdx.x;
    DD acc = {0.0, 0.0};
    for (int64_t b = t; b < n; b += 256) {
        DD x = {partial[2 * b], partial[2 * b + 1]};
        acc = dd_add(acc, x);
    }
    ss[t] = acc.s;
    sc[t] = acc.c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            DD a = {ss[t], sc[t]}, b = {ss[t + w], sc[t + w]};
            DD r = dd_add(a, b);
            ss[t] = r.s;
            sc[t] = r.c;
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = ss[0];
        out[1] = sc[0];
    }
}

// Rounding residual of an FP32 state: the tiles' running totals summed in a fixed order
// (strided over the work-items, then a tree) into out (T sum, T compensation, S sum,
// S compensation).
constexpr int RB = 1024;   // work-items of the reduction
__global__ void __launch_bounds__(RB) residual_reduce_kernel(const double* residual, int nslots, double* out) {
    __shared__ double ts[RB], tc[RB], ss[RB], sc[RB];
    const int t = threadIdx.x;
    const double4* slot = reinterpret_cast<const double4*>(residual);
    DD aT = {0.0, 0.0}, aS = {0.0, 0.0};
    for (int b = t; b < nslots; b += RB) {
        const double4 v = slot[b];
        aT = dd_add(aT, DD{v.x, v.y});
        aS = dd_add(aS, DD{v.z, v.w});
    }
    ts[t] = aT.s;
    tc[t] = aT.c;
    ss[t] = aS.s;
    sc[t] = aS.c;
    __syncthreads();
    for (int w = RB / 2; w > 0; w >>= 1) {
        if (t < w) {
            const DD rT = dd_add(DD{ts[t], tc[t]}, DD{ts[t + w], tc[t + w]});
            const DD rS = dd_add(DD{ss[t], sc[t]}, DD{ss[t + w], sc[t + w]});
            ts[t] = rT.s;
            tc[t] = rT.c;
            ss[t] = rS.s;
            sc[t] = rS.c;
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = ts[0];
        out[1] = tc[0];
        out[2] = ss[0];
        out[3] = sc[0];
    }
}

// Traffic probe: out_j[i] = Σ_k in_k[i] + j over n doubles, i fastest (one double per
// work-item, grid-stride), so a launch moves exactly (n_in + n_out)·n·8 bytes of HBM.
struct StreamProbeArgs {
    const double* in[NE_STREAM_PROBE_MAX_IN];
    double* out[NE_STREAM_PROBE_MAX_OUT];
    int n_in, n_out;
    int64_t n;
};

__global__ void __launch_bounds__(256) stream_probe_kernel(StreamProbeArgs a) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * 256) {
        double s = 0.0;
        for (int k = 0; k < a.n_in; k++) s += a.in[k][i];
        for (int j = 0; j < a.n_out; j++) a.out[j][i] = s + (double)j;
    }
}

// Halo of a cell-centred field: every halo point takes the value of the interior point it maps
// to — wrapped in a periodic direction, the nearest interior point in a bounded one and
// vertically (the zero-gradient halo of a flux boundary condition). Interior points are only
// read and each halo point is written once, so the work-items are independent.
template <typename Real>
__global__ void __launch_bounds__(256) halo_fill_kernel(View3<Real> c, int Nx, int Ny, int Nz, int Hx, int Hy,
                                                        int Hz, int periodic_x, int periodic_y) {
    const int64_t W = Nx + 2 * Hx, H = Ny + 2 * Hy, n = W * H * (int64_t)(Nz + 2 * Hz);
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int i = (int)(t % W) - Hx;
        const int j = (int)((t / W) % H) - Hy;
        const int k = (int)(t / (W * H)) - Hz;
        if (i >= 0 && i < Nx && j >= 0 && j < Ny && k >= 0 && k < Nz) continue;
        const int si = periodic_x ? ((i % Nx) + Nx) % Nx : min(max(i, 0), Nx - 1);
        const int sj = periodic_y ? ((j % Ny) + Ny) % Ny : min(max(j, 0), Ny - 1);
        const int sk = min(max(k, 0), Nz - 1);
        c(i, j, k) = c(si, sj, sk);
    }
}

// Beer's-law factors at the faces, with the expressions the stage kernel evaluates in place.
__global__ void beer_factors_kernel(const double* __restrict__ zf, int Nz, double kappa1, double kappa2,
                                    double* __restrict__ beer) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k > Nz) return;
    const double eta = zf[Nz];
    beer[2 * k] = exp(kappa1 * (zf[k] - eta));
    beer[2 * k + 1] = exp(kappa2 * (zf[k] - eta));
}

}  // namespace

hipError_t launch_beer_factors(const GridDev& g, double kappa1, double kappa2, double* beer, hipStream_t s) {
    hipLaunchKernelGGL(beer_factors_kernel, dim3((g.Nz + 1 + 63) / 64), dim3(64), 0, s, g.zf, g.Nz, kappa1, kappa2,
                       beer);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_halo_fill(const GridDev& g, View3<Real> c, int Hx, int Hy, int Hz, hipStream_t s) {
    const int64_t n = (int64_t)(g.Nx + 2 * Hx) * (g.Ny + 2 * Hy) * (g.Nz + 2 * Hz);
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(halo_fill_kernel<Real>, dim3((unsigned)blocks), dim3(256), 0, s, c, g.Nx, g.Ny, g.Nz, Hx, Hy, Hz,
                       g.tx == NE_PERIODIC ? 1 : 0, g.ty == NE_PERIODIC ? 1 : 0);
    return hipGetLastError();
}
template hipError_t launch_halo_fill<double>(const GridDev&, View3<double>, int, int, int, hipStream_t);
template hipError_t launch_halo_fill<float>(const GridDev&, View3<float>, int, int, int, hipStream_t);

hipError_t launch_stream_probe(const double* const* in, int n_in, double* const* out, int n_out, int64_t n,
                               hipStream_t s) {
    StreamProbeArgs a = {};
    for (int k = 0; k < n_in; k++) a.in[k] = in[k];
    for (int j = 0; j < n_out; j++) a.out[j] = out[j];
    a.n_in = n_in;
    a.n_out = n_out;
    a.n = n;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(stream_probe_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename Real>
TracerLayout host_layout(const View3<const Real>& c, const View3<const Real>& u, const View3<const Real>& v,
                         const View3<const Real>& w, const View3<const double>& J) {
    TracerLayout L;
    L.c_sy = (uint32_t)(c.sy * (int64_t)sizeof(Real));
    L.c_sz = (uint32_t)(c.sz * (int64_t)sizeof(Real));
    L.u_sy = (uint32_t)(u.sy * (int64_t)sizeof(Real));
    L.u_sz = (uint32_t)(u.sz * (int64_t)sizeof(Real));
    L.v_sy = (uint32_t)(v.sy * (int64_t)sizeof(Real));
    L.v_sz = (uint32_t)(v.sz * (int64_t)sizeof(Real));
    L.w_sy = (uint32_t)(w.sy * (int64_t)sizeof(Real));
    L.w_sz = (uint32_t)(w.sz * (int64_t)sizeof(Real));
    L.j_sy = (uint32_t)(J.sy * (int64_t)sizeof(double));
    return L;
}

template <typename Real>
hipError_t launch_rk_stage(const GridDev& g, View3<const Real> u, View3<const Real> v, View3<const Real> w,
                           View3<const Real> T0, View3<const Real> S0, View3<const Real> Tin, View3<const Real> Sin,
                           View3<Real> Tout, View3<Real> Sout, View3<const double> JT, View3<const double> JS,
                           View3<const double> Jw, TracerForcing forcing, double beta_dt, bool stochastic,
                           double time, hipStream_t s) {
    TracerLayout L = host_layout<Real>(Tin, u, v, w, JT);
    const dim3 grid(grid_blocks(g.Nx, g.Ny)), block(TW, PLAIN_TR);
    const int sr = stochastic ? 1 : 0;
    const uint32_t seed = step_seed(time);
    const bool watch = forcing.freeze_flags != nullptr;
    const bool beer = forcing.beer != nullptr;   // a model's factors (per-phase calls compute them)
    if (forcing.two_color && watch && beer)
        hipLaunchKernelGGL((tracer_tile_kernel<Real, 0, true, true, true>), grid, block, 0, s, g, L, u.p, v.p, w.p,
                           T0.p, S0.p, Tin.p, Sin.p, Tout.p, Sout.p, JT.p, JS.p, Jw.p, forcing, beta_dt, sr, seed);
    else if (forcing.two_color && beer)
        hipLaunchKernelGGL((tracer_tile_kernel<Real, 0, true, false, true>), grid, block, 0, s, g, L, u.p, v.p, w.p,
                           T0.p, S0.p, Tin.p, Sin.p, Tout.p, Sout.p, JT.p, JS.p, Jw.p, forcing, beta_dt, sr, seed);
    else if (forcing.two_color && watch)
        hipLaunchKernelGGL((tracer_tile_kernel<Real, 0, true, true>), grid, block, 0, s, g, L, u.p, v.p, w.p, T0.p,
                           S0.p, Tin.p, Sin.p, Tout.p, Sout.p, JT.p, JS.p, Jw.p, forcing, beta_dt, sr, seed);
    else if (forcing.two_color)
        hipLaunchKernelGGL((tracer_tile_kernel<Real, 0, true, false>), grid, block, 0, s, g, L, u.p, v.p, w.p, T0.p,
                           S0.p, Tin.p, Sin.p, Tout.p, Sout.p, JT.p, JS.p, Jw.p, forcing, beta_dt, sr, seed);
    else if (watch)
        hipLaunchKernelGGL((tracer_tile_kernel<Real, 0, false, true>), grid, block, 0, s, g, L, u.p, v.p, w.p, T0.p,
                           S0.p, Tin.p, Sin.p, Tout.p, Sout.p, JT.p, JS.p, Jw.p, forcing, beta_dt, sr, seed);
    else
        hipLaunchKernelGGL((tracer_tile_kernel<Real, 0, false, false>), grid, block, 0, s, g, L, u.p, v.p, w.p, T0.p,
                           S0.p, Tin.p, Sin.p, Tout.p, Sout.p, JT.p, JS.p, Jw.p, forcing, beta_dt, sr, seed);
    return hipGetLastError();
}
template hipError_t launch_rk_stage<double>(const GridDev&, View3<const double>, View3<const double>,
                                            View3<const double>, View3<const double>, View3<const double>,
                                            View3<const double>, View3<const double>, View3<double>, View3<double>,
                                            View3<const double>, View3<const double>, View3<const double>,
                                            TracerForcing, double, bool, double, hipStream_t);
template hipError_t launch_rk_stage<float>(const GridDev&, View3<const float>, View3<const float>,
                                           View3<const float>, View3<const float>, View3<const float>,
                                           View3<const float>, View3<const float>, View3<float>, View3<float>,
                                           View3<const double>, View3<const double>, View3<const double>,
                                           TracerForcing, double, bool, double, hipStream_t);

// Quasi-Adams–Bashforth 2 update (the reference's timestepper = :QuasiAdamsBashforth2,
// ocean_simulation.jl:323, with Oceananigans' default χ = 0.1): in every active cell
// c ← c + Δt·((3/2 + χ)·Gⁿ − (1/2 + χ)·G⁻), one work-item per column taking the levels in turn
// (i fastest across a wavefront). A step taken as forward Euler (the first step, or a Δt other
// than the last step's) has χ = −1/2 and reads no G⁻ (Oceananigans zeroes it). The column's flag
// is set when a stored cell is below its freezing point, the test the RK step's last stage makes.
__global__ void __launch_bounds__(256) ab2_update_kernel(GridDev g, View3<double> T, View3<double> S,
                                                        View3<const double> GT, View3<const double> GS,
                                                        View3<const double> GmT, View3<const double> GmS,
                                                        double dt, double a, double b, int euler,
                                                        uint8_t* freeze_flags, double melt0, double slope) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    if (i >= g.Nx || j >= g.Ny) return;
    const int64_t col = (int64_t)i + (int64_t)j * g.Nx;
    bool frz = false;
    for (int k = g.kmin[col]; k < g.Nz; k++) {
        const double gmT = euler ? 0.0 : GmT(i, j, k), gmS = euler ? 0.0 : GmS(i, j, k);
        const double Tn = T(i, j, k) + dt * (a * GT(i, j, k) - b * gmT);
        const double Sn = S(i, j, k) + dt * (a * GS(i, j, k) - b * gmS);
        T(i, j, k) = Tn;
        S(i, j, k) = Sn;
        frz = frz || Tn < melt0 - slope * Sn;
    }
    if (freeze_flags) freeze_flags[col] = frz ? 1 : 0;
}

hipError_t launch_ab2_update(const GridDev& g, View3<double> T, View3<double> S, View3<const double> GT,
                             View3<const double> GS, View3<const double> GmT, View3<const double> GmS, double dt,
                             double chi, bool euler, uint8_t* freeze_flags, double melt0, double slope,
                             hipStream_t s) {
    const double c = euler ? -0.5 : chi;
    hipLaunchKernelGGL(ab2_update_kernel, dim3((g.Nx + 63) / 64, (g.Ny + 3) / 4), dim3(64, 4), 0, s, g, T, S, GT, GS,
                       GmT, GmS, dt, 1.5 + c, 0.5 + c, euler ? 1 : 0, freeze_flags, melt0, slope);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_tendency(const GridDev& g, View3<const Real> u, View3<const Real> v, View3<const Real> w,
                           View3<const Real> T, View3<const Real> S, View3<const double> JT,
                           View3<const double> JS, View3<const double> Jw, TracerForcing forcing, View3<Real> GT,
                           View3<Real> GS, hipStream_t s) {
    TracerLayout L = host_layout<Real>(T, u, v, w, JT);
    if (forcing.two_color)
        hipLaunchKernelGGL((tracer_tile_kernel<Real, 1, true>), dim3(grid_blocks(g.Nx, g.Ny)), dim3(TW, PLAIN_TR), 0,
                           s, g, L, u.p, v.p, w.p, T.p, S.p, T.p, S.p, GT.p, GS.p, JT.p, JS.p, Jw.p, forcing, 0.0, 0,
                           0u);
    else
        hipLaunchKernelGGL((tracer_tile_kernel<Real, 1, false>), dim3(grid_blocks(g.Nx, g.Ny)), dim3(TW, PLAIN_TR),
                           0, s, g, L, u.p, v.p, w.p, T.p, S.p, T.p, S.p, GT.p, GS.p, JT.p, JS.p, Jw.p, forcing, 0.0,
                           0, 0u);
    return hipGetLastError();
}
template hipError_t launch_tendency<double>(const GridDev&, View3<const double>, View3<const double>,
                                            View3<const double>, View3<const double>, View3<const double>,
                                            View3<const double>, View3<const double>, View3<const double>,
                                            TracerForcing, View3<double>, View3<double>, hipStream_t);
template hipError_t launch_tendency<float>(const GridDev&, View3<const float>, View3<const float>,
                                           View3<const float>, View3<const float>, View3<const float>,
                                           View3<const double>, View3<const double>, View3<const double>,
                                           TracerForcing, View3<float>, View3<float>, hipStream_t);

template <typename Real>
hipError_t launch_budget(ne_ctx* ctx, const GridDev& g, View3<const Real> c, double weight, double* out,
                         hipStream_t s) {
    dim3 gr = grid2d(g.Nx, g.Ny);
    int64_t nb = (int64_t)gr.x * gr.y;
    size_t need = (size_t)nb * 2 * sizeof(double);
    if (ctx->scratch_bytes < need) {
        if (ctx->scratch) {
            hipError_t e = hipFree(ctx->scratch);
            if (e != hipSuccess) return e;
            ctx->scratch = nullptr;
            ctx->scratch_bytes = 0;
        }
        hipError_t e = hipMalloc((void**)&ctx->scratch, need);
        if (e != hipSuccess) return e;
        ctx->scratch_bytes = need;
    }
    hipLaunchKernelGGL(budget_partial_kernel<Real>, gr, dim3(BX, BY), 0, s, g, c, weight, ctx->scratch);
    hipLaunchKernelGGL(budget_final_kernel, dim3(1), dim3(256), 0, s, ctx->scratch, nb, out);
    return hipGetLastError();
}
template hipError_t launch_budget<double>(ne_ctx*, const GridDev&, View3<const double>, double, double*,
                                          hipStream_t);

namespace {
std::mutex noalloc_mu;
std::vector<double*> noalloc_ptr;   // per device: the address of `scratch` (nullptr: not looked up yet)
}  // namespace

hipError_t budget_noalloc_init(int device) {
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if ((int)noalloc_ptr.size() <= device) noalloc_ptr.resize(device + 1, nullptr);
    if (noalloc_ptr[device]) return hipSuccess;
    int current = -1;
    hipError_t e = hipGetDevice(&current);
    if (e != hipSuccess) return e;
    if (current != device) return hipErrorInvalidDevice;   // the caller makes `device` current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    noalloc_ptr[device] = static_cast<double*>(p);
    return hipSuccess;
}

double* budget_noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lock(noalloc_mu);
    return device >= 0 && device < (int)noalloc_ptr.size() ? noalloc_ptr[device] : nullptr;
}

int64_t budget_blocks(const GridDev& g) {
    const dim3 gr = grid2d(g.Nx, g.Ny);
    return (int64_t)gr.x * gr.y;
}

// The kernels of launch_budget with the partials in the device's `scratch` array; allocates
// nothing. Every block overwrites its own pair, so nothing needs clearing between calls.
// One slot of 4 doubles per tile of the tiling with the most tiles (the fewest rows: the
// single-model tiling or the batch's 4-row one); a batch of taller tiles uses the first slots.
int64_t residual_slots(const GridDev& g) {
    return (int64_t)tiles_x(g.Nx) * tiles_y(g.Ny, PLAIN_TR < 4 ? PLAIN_TR : 4);
}

hipError_t launch_residual_reduce(const double* residual, const GridDev& g, double* out, hipStream_t s) {
    hipLaunchKernelGGL(residual_reduce_kernel, dim3(1), dim3(RB), 0, s, residual, (int)residual_slots(g), out);
    return hipGetLastError();
}

hipError_t launch_budget_noalloc(double* partial, const GridDev& g, View3<const double> c, double weight,
                                 double* out, hipStream_t s) {
    const dim3 gr = grid2d(g.Nx, g.Ny);
    const int64_t nb = (int64_t)gr.x * gr.y;
    hipLaunchKernelGGL(budget_partial_kernel<double>, gr, dim3(BX, BY), 0, s, g, c, weight, partial);
    hipLaunchKernelGGL(budget_final_kernel, dim3(1), dim3(256), 0, s, partial, nb, out);
    return hipGetLastError();
}
template hipError_t launch_budget<float>(ne_ctx*, const GridDev&, View3<const float>, double, double*,
                                         hipStream_t);

template <typename Real>
hipError_t launch_w_continuity(const GridDev& g, View3<const Real> u, View3<const Real> v, View3<Real> w,
                               hipStream_t s) {
    hipLaunchKernelGGL(w_continuity_kernel<Real>, dim3((g.Nx + 63) / 64, (g.Ny + 3) / 4), dim3(64, 4), 0, s, g, u, v,
                       w);
    return hipGetLastError();
}
template hipError_t launch_w_continuity<double>(const GridDev&, View3<const double>, View3<const double>,
                                                View3<double>, hipStream_t);
template hipError_t launch_w_continuity<float>(const GridDev&, View3<const float>, View3<const float>, View3<float>,
                                               hipStream_t);

int batch_tiles_x(int Nx) { return tiles_x(Nx); }
int batch_tiles_y(int Ny, int tile_rows) { return tiles_y(Ny, tile_rows); }

template <typename Real, bool TWO_COLOR, bool WATCH>
hipError_t launch_batch_stage_w(const BatchRec<Real>* recs, const WorkItem* work, int nwork, int tile_rows,
                                int blocks, int stage, double beta_dt, hipStream_t s) {
    switch (tile_rows) {
    case 4:
        hipLaunchKernelGGL((batch_stage_kernel<Real, 4, TWO_COLOR, WATCH>), dim3(blocks), dim3(TW, 4), 0, s, recs,
                           work, nwork, stage, beta_dt);
        break;
    case 8:
        hipLaunchKernelGGL((batch_stage_kernel<Real, 8, TWO_COLOR, WATCH>), dim3(blocks), dim3(TW, 8), 0, s, recs,
                           work, nwork, stage, beta_dt);
        break;
    case 16:
        hipLaunchKernelGGL((batch_stage_kernel<Real, 16, TWO_COLOR, WATCH>), dim3(blocks), dim3(TW, 16), 0, s, recs,
                           work, nwork, stage, beta_dt);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Stage 3 (the step's last) is the instance that flags the columns the freeze clamp visits.
template <typename Real, bool TWO_COLOR>
hipError_t launch_batch_stage_tc(const BatchRec<Real>* recs, const WorkItem* work, int nwork, int tile_rows,
                                 int blocks, int stage, double beta_dt, hipStream_t s) {
    return stage == 3 ? launch_batch_stage_w<Real, TWO_COLOR, true>(recs, work, nwork, tile_rows, blocks, stage,
                                                                    beta_dt, s)
                      : launch_batch_stage_w<Real, TWO_COLOR, false>(recs, work, nwork, tile_rows, blocks, stage,
                                                                     beta_dt, s);
}

template <typename Real>
hipError_t launch_batch_stage(const BatchRec<Real>* recs, const WorkItem* work, int nwork, int tile_rows,
                              int blocks, int stage, double beta_dt, bool two_color, hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    if (blocks <= 0 || blocks % NXCD != 0) return hipErrorInvalidValue;
    return two_color ? launch_batch_stage_tc<Real, true>(recs, work, nwork, tile_rows, blocks, stage, beta_dt, s)
                     : launch_batch_stage_tc<Real, false>(recs, work, nwork, tile_rows, blocks, stage, beta_dt, s);
}
template hipError_t launch_batch_stage<double>(const BatchRec<double>*, const WorkItem*, int, int, int, int, double,
                                               bool, hipStream_t);
template hipError_t launch_batch_stage<float>(const BatchRec<float>*, const WorkItem*, int, int, int, int, double,
                                              bool, hipStream_t);

}  // namespace ne
