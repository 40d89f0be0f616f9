// ne_tracer.hip — WENO5 tracer-advection tendency on the immersed LatLon grid, fused
// with the top flux boundary condition and the split Runge-Kutta stage update.
//
// Tiling (gfx950): a block of 512 work-items = 8 wavefronts owns a tile of 63 × 8
// columns and marches up the active levels. At each level the tile's T and S plane
// plus a 3-cell halo (69 × 14 cells) is staged in LDS; every x- and y-face flux of the
// tile is reconstructed once (lane ℓ of wavefront r owns the west face of column ℓ and
// the south face of row r, lane 63 only a face), written to LDS, and differenced by the
// owning cell. The vertical face flux at the top of level k is carried in registers as
// the bottom flux of level k+1, with the column's vertical stencil in a register window.
// Tiles are ordered so that each XCD (blocks are dealt round-robin to the 8 XCDs)
// receives a contiguous band of tiles, keeping halo re-reads inside one L2.
#include "ne_internal.h"

namespace ne {

namespace {

constexpr int TW = 64;            // work-items per tile row (x faces per row)
constexpr int TC = 63;            // cells per tile row
constexpr int PLAIN_TR = 8;       // cell rows per tile of the single-model kernel
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
};

__host__ __device__ inline int tiles_x(int Nx) { return (Nx + TC - 1) / TC; }
__host__ __device__ inline int tiles_y(int Ny, int tr) { return (Ny + tr - 1) / tr; }
inline int grid_blocks(int Nx, int Ny) {
    int n = tiles_x(Nx) * tiles_y(Ny, PLAIN_TR);
    return ((n + NXCD - 1) / NXCD) * NXCD;
}

// Activity bits of six consecutive cells whose kmin is at km[0], km[stride], …
__device__ __forceinline__ unsigned act6(const int* km, int stride, int k) {
    unsigned a = 0;
#pragma unroll
    for (int n = 0; n < 6; n++) a |= (k >= km[n * stride] ? 1u : 0u) << n;
    return a;
}

// One tile (columns i0 … i0+62, rows j0 … j0+TR-1) marched over its active levels.
// MODE 0: c_out = c0 + beta_dt·G(c_in) on active cells. MODE 1: write G (0 on inactive
// cells). Every work-item of the block calls this with the same tile.
template <typename Real, int MODE, int TR>
__device__ __forceinline__ void tile_body(TileShared<Real, TR>& sh, const GridDev& g, const View3<const Real>& u,
                                          const View3<const Real>& v, const View3<const Real>& w,
                                          const View3<const Real>& T0, const View3<const Real>& S0,
                                          const View3<const Real>& Tin, const View3<const Real>& Sin,
                                          const View3<Real>& Tout, const View3<Real>& Sout,
                                          const View3<const double>& JT, const View3<const double>& JS,
                                          const View3<const double>& Jw, double beta_dt, int i0, int j0) {
    typedef Real R;
    constexpr int PXW = TileShared<R, TR>::PXW;
    constexpr int PN = TileShared<R, TR>::PN;
    constexpr int NTHR = TW * TR;
    const int lane = threadIdx.x;
    const int row = threadIdx.y;
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
    __syncthreads();
    const int kstart = sh.kstart;
    if (MODE == 1 && cell)
        for (int k = 0; k < mykmin && k < Nz; k++) { Tout(ci, cj, k) = (R)0; Sout(ci, cj, k) = (R)0; }
    if (kstart >= Nz) return;  // uniform: the whole tile is land

    // Per-row and per-column geometry
    const bool xrow = cj < g.Ny;                 // this wavefront's row is inside the domain
    const int jw = xrow ? cj : 0;
    const double Axf = g.dy_xface[jw];
    const int fiw = g.wx(i0 + lane);             // x face owned by this lane (west face of column i0+lane)
    const bool ycol = lane < TC && ci < g.Nx;
    const int cy = ycol ? ci : 0;
    const int fjs = g.wy(j0 + row);              // y face owned by this lane (south face of row j0+row)
    const int fjn = g.wy(j0 + TR);               // north face of the last row (wavefront TR-1)
    const double Ays = (j0 + row <= g.Ny) ? g.dx_yface[fjs < 0 ? 0 : (fjs > g.Ny ? g.Ny : fjs)] : 0.0;
    const double Ayn = (j0 + TR <= g.Ny) ? g.dx_yface[fjn < 0 ? 0 : (fjn > g.Ny ? g.Ny : fjn)] : 0.0;
    const double Azc = cell ? g.Az[cj] : 0.0;

    // Vertical register window of the owned column: levels k-2 … k+3
    R zT[6], zS[6];
#pragma unroll
    for (int n = 0; n < 6; n++) {
        int kk = kstart - 2 + n;
        bool a = cell && kk >= mykmin && kk < Nz;
        zT[n] = a ? Tin(ci, cj, kk) : (R)0;
        zS[n] = a ? Sin(ci, cj, kk) : (R)0;
    }
    R FzT = (R)0, FzS = (R)0;  // flux through the bottom face of the current level

    for (int k = kstart; k < Nz; k++) {
        // 1. stage the plane of level k
        for (int n = tid; n < PN; n += NTHR) {
            bool a = k >= sh.km[n];
            R t = (R)0, s = (R)0;
            if (a) {
                int gi = g.wx(i0 - 3 + n % PXW), gj = g.wy(j0 - 3 + n / PXW);
                t = Tin(gi, gj, k);
                s = Sin(gi, gj, k);
            }
            sh.cT[n] = t;
            sh.cS[n] = s;
        }
        __syncthreads();
        const R dzk = (R)g.dz[k];

        // 2a. x face owned by this lane: cells i0+lane-3 … i0+lane+2 of row j0+row
        {
            R fT = (R)0, fS = (R)0;
            if (xrow) {
                const int base = (row + 3) * PXW + lane;
                int order = adv ? face_order(act6(&sh.km[base], 1, k)) : 0;
                if (order > 0) {
                    R vel = u(fiw, cj, k);
                    R A = (R)Axf * dzk;
                    const R* t = &sh.cT[base];
                    const R* s = &sh.cS[base];
                    fT = A * vel * face_value(t[0], t[1], t[2], t[3], t[4], t[5], order, vel);
                    fS = A * vel * face_value(s[0], s[1], s[2], s[3], s[4], s[5], order, vel);
                }
            }
            sh.fxT[row * TW + lane] = fT;
            sh.fxS[row * TW + lane] = fS;
        }
        // 2b. y faces: the south face of row j0+row, and for the last wavefront the north face of the tile
        if (ycol) {
#pragma unroll 1
            for (int pass = 0; pass < (row == TR - 1 ? 2 : 1); pass++) {
                const int r0 = row + pass;                // plane row of the first stencil cell
                const int base = r0 * PXW + lane + 3;
                const int fj = pass ? fjn : fjs;
                const double Ay = pass ? Ayn : Ays;
                R fT = (R)0, fS = (R)0;
                int order = adv ? face_order(act6(&sh.km[base], PXW, k)) : 0;
                if (order > 0) {
                    R vel = v(cy, fj, k);
                    R A = (R)Ay * dzk;
                    const R* t = &sh.cT[base];
                    const R* s = &sh.cS[base];
                    fT = A * vel * face_value(t[0], t[PXW], t[2 * PXW], t[3 * PXW], t[4 * PXW], t[5 * PXW], order, vel);
                    fS = A * vel * face_value(s[0], s[PXW], s[2 * PXW], s[3 * PXW], s[4 * PXW], s[5 * PXW], order, vel);
                }
                sh.fyT[(r0)*TC + lane] = fT;
                sh.fyS[(r0)*TC + lane] = fS;
            }
        }
        // 2c. z face at the top of level k (cells k-2 … k+3 of the owned column)
        R FztT = (R)0, FztS = (R)0;
        if (cell && adv) {
            unsigned az = 0;
#pragma unroll
            for (int n = 0; n < 6; n++) {
                int kk = k - 2 + n;
                az |= ((kk >= mykmin && kk < Nz) ? 1u : 0u) << n;
            }
            int order = face_order(az);
            if (order > 0) {
                R vel = w(ci, cj, k + 1);
                R A = (R)Azc;
                FztT = A * vel * face_value(zT[0], zT[1], zT[2], zT[3], zT[4], zT[5], order, vel);
                FztS = A * vel * face_value(zS[0], zS[1], zS[2], zS[3], zS[4], zS[5], order, vel);
            }
        }
        __syncthreads();

        // 3. difference the face fluxes, add the top boundary flux, update
        if (cell && k >= mykmin) {
            const int xo = row * TW + lane, yo = row * TC + lane;
            R div_T = (sh.fxT[xo + 1] - sh.fxT[xo]) + (sh.fyT[yo + TC] - sh.fyT[yo]);
            div_T = div_T + (FztT - FzT);
            R div_S = (sh.fxS[xo + 1] - sh.fxS[xo]) + (sh.fyS[yo + TC] - sh.fyS[yo]);
            div_S = div_S + (FztS - FzS);
            const R V = (R)(Azc * g.dz[k]);
            R GT = -((R)1 / V * div_T);
            R GS = -((R)1 / V * div_S);
            if (k == Nz - 1) {
                // MultipleFluxes(Jᵀ, FreshwaterExchange{:T}) = Jᵀ + 0 on a static grid;
                // MultipleFluxes(Jˢ, FreshwaterExchange{:S}) = Jˢ + (Sᴺ Jʷ − 0)
                // (multiple_surface_fluxes.jl:27-31; ocean_simulation.jl:80-91)
                double FT = JT.at2(ci, cj) + (0.0 + 0.0);
                double FS = JS.at2(ci, cj) + (((double)zS[2] * Jw.at2(ci, cj) - 0.0) + 0.0);
                GT = GT - (R)(FT * Azc) / V;
                GS = GS - (R)(FS * Azc) / V;
            }
            if (MODE == 0) {
                // The update is formed in FP64 and rounded once to the storage type (for FP32
                // storage this keeps the per-stage rounding of the state at half an ulp).
                Tout(ci, cj, k) = (R)((double)T0(ci, cj, k) + beta_dt * (double)GT);
                Sout(ci, cj, k) = (R)((double)S0(ci, cj, k) + beta_dt * (double)GS);
            } else {
                Tout(ci, cj, k) = GT;
                Sout(ci, cj, k) = GS;
            }
        }
        FzT = FztT;
        FzS = FztS;
        // 4. slide the vertical window up one level
        if (cell) {
#pragma unroll
            for (int n = 0; n < 5; n++) { zT[n] = zT[n + 1]; zS[n] = zS[n + 1]; }
            int kk = k + 4;
            bool a = kk >= mykmin && kk < Nz;
            zT[5] = a ? Tin(ci, cj, kk) : (R)0;
            zS[5] = a ? Sin(ci, cj, kk) : (R)0;
        }
    }
}

// Single model: one block per tile, tiles ordered so that each XCD (blocks are dealt
// round-robin to the 8 XCDs) receives a contiguous band of tiles.
template <typename Real, int MODE>
__global__ void __launch_bounds__(TW * PLAIN_TR)
tracer_tile_kernel(GridDev g, View3<const Real> u, View3<const Real> v, View3<const Real> w, View3<const Real> T0,
                   View3<const Real> S0, View3<const Real> Tin, View3<const Real> Sin, View3<Real> Tout,
                   View3<Real> Sout, View3<const double> JT, View3<const double> JS, View3<const double> Jw,
                   double beta_dt) {
    __shared__ TileShared<Real, PLAIN_TR> sh;
    const int ntx = tiles_x(g.Nx), nty = tiles_y(g.Ny, PLAIN_TR);
    const int ntiles = ntx * nty;
    const int per = (ntiles + NXCD - 1) / NXCD;
    const int tile = (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
    if (tile >= ntiles) return;
    tile_body<Real, MODE, PLAIN_TR>(sh, g, u, v, w, T0, S0, Tin, Sin, Tout, Sout, JT, JS, Jw, beta_dt,
                                    (tile % ntx) * TC, (tile / ntx) * PLAIN_TR);
}

// Ensemble: a persistent grid walks the (member, tile) work list. The list is cut into 8
// contiguous chunks, chunk x going to the blocks of XCD x, which stride through it; each
// member's tiles are contiguous in the list, so a member's halo re-reads stay in one L2.
// stage 1: T¹ = T + β₁Δt G(T); stage 2: T² = T + β₂Δt G(T¹); stage 3: T = T + β₃Δt G(T²).
template <typename Real, int TR>
__global__ void __launch_bounds__(TW * TR)
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
        const int ntx = tiles_x(g.Nx);
        const View3<const Real> Tin = cview(stage == 1 ? rc.T : stage == 2 ? rc.T1 : rc.T2);
        const View3<const Real> Sin = cview(stage == 1 ? rc.S : stage == 2 ? rc.S1 : rc.S2);
        const View3<Real> Tout = stage == 1 ? rc.T1 : stage == 2 ? rc.T2 : rc.T;
        const View3<Real> Sout = stage == 1 ? rc.S1 : stage == 2 ? rc.S2 : rc.S;
        __syncthreads();   // the previous item's reads of the LDS are complete
        tile_body<Real, 0, TR>(sh, g, rc.u, rc.v, rc.w, cview(rc.T), cview(rc.S), Tin, Sin, Tout, Sout,
                               cview(rc.n.J_T), cview(rc.n.J_S), cview(rc.n.J_w), beta_dt,
                               (it.index % ntx) * TC, (it.index / ntx) * TR);
    }
}

// ---------------------------------------------------------------------------
// Compensated budget: Neumaier per column, then a fixed-order double-double tree.
// ---------------------------------------------------------------------------

constexpr int BX = 64;
constexpr int BY = 4;

inline dim3 grid2d(int Nx, int Ny) { return dim3((Nx + BX - 1) / BX, (Ny + BY - 1) / BY); }

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

}  // namespace

template <typename Real>
hipError_t launch_rk_stage(const GridDev& g, View3<const Real> u, View3<const Real> v, View3<const Real> w,
                           View3<const Real> T0, View3<const Real> S0, View3<const Real> Tin, View3<const Real> Sin,
                           View3<Real> Tout, View3<Real> Sout, View3<const double> JT, View3<const double> JS,
                           View3<const double> Jw, double beta_dt, hipStream_t s) {
    hipLaunchKernelGGL((tracer_tile_kernel<Real, 0>), dim3(grid_blocks(g.Nx, g.Ny)), dim3(TW, PLAIN_TR), 0, s, g, u, v, w,
                       T0, S0, Tin, Sin, Tout, Sout, JT, JS, Jw, beta_dt);
    return hipGetLastError();
}
template hipError_t launch_rk_stage<double>(const GridDev&, View3<const double>, View3<const double>,
                                            View3<const double>, View3<const double>, View3<const double>,
                                            View3<const double>, View3<const double>, View3<double>, View3<double>,
                                            View3<const double>, View3<const double>, View3<const double>, double,
                                            hipStream_t);
template hipError_t launch_rk_stage<float>(const GridDev&, View3<const float>, View3<const float>,
                                           View3<const float>, View3<const float>, View3<const float>,
                                           View3<const float>, View3<const float>, View3<float>, View3<float>,
                                           View3<const double>, View3<const double>, View3<const double>, double,
                                           hipStream_t);

hipError_t launch_tendency(const GridDev& g, View3<const double> u, View3<const double> v, View3<const double> w,
                           View3<const double> T, View3<const double> S, View3<const double> JT,
                           View3<const double> JS, View3<const double> Jw, View3<double> GT, View3<double> GS,
                           hipStream_t s) {
    hipLaunchKernelGGL((tracer_tile_kernel<double, 1>), dim3(grid_blocks(g.Nx, g.Ny)), dim3(TW, PLAIN_TR), 0, s, g, u, v,
                       w, T, S, T, S, GT, GS, JT, JS, Jw, 0.0);
    return hipGetLastError();
}

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
template hipError_t launch_budget<float>(ne_ctx*, const GridDev&, View3<const float>, double, double*,
                                         hipStream_t);

int batch_tiles_x(int Nx) { return tiles_x(Nx); }
int batch_tiles_y(int Ny, int tile_rows) { return tiles_y(Ny, tile_rows); }

template <typename Real>
hipError_t launch_batch_stage(const BatchRec<Real>* recs, const WorkItem* work, int nwork, int tile_rows,
                              int blocks, int stage, double beta_dt, hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    if (blocks <= 0 || blocks % NXCD != 0) return hipErrorInvalidValue;
    switch (tile_rows) {
    case 4:
        hipLaunchKernelGGL((batch_stage_kernel<Real, 4>), dim3(blocks), dim3(TW, 4), 0, s, recs, work, nwork, stage,
                           beta_dt);
        break;
    case 8:
        hipLaunchKernelGGL((batch_stage_kernel<Real, 8>), dim3(blocks), dim3(TW, 8), 0, s, recs, work, nwork, stage,
                           beta_dt);
        break;
    case 16:
        hipLaunchKernelGGL((batch_stage_kernel<Real, 16>), dim3(blocks), dim3(TW, 16), 0, s, recs, work, nwork,
                           stage, beta_dt);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_batch_stage<double>(const BatchRec<double>*, const WorkItem*, int, int, int, int, double,
                                               hipStream_t);
template hipError_t launch_batch_stage<float>(const BatchRec<float>*, const WorkItem*, int, int, int, int, double,
                                              hipStream_t);

}  // namespace ne
