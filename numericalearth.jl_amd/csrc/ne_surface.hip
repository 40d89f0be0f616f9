// ne_surface.hip — surface phases of the coupled step on gfx950.
//
// Layout: one work-item per ocean column (i, j); blocks of 64×4 with i fastest so
// that each wavefront reads 64 consecutive columns of every 2-D field (coalesced).
// The flux solve is FP64-VALU bound (transcendental-heavy fixed point); lanes of a
// wavefront retire independently and the wavefront leaves the loop when its last
// lane has converged (the loop condition is the wavefront-wide vote).
#include <type_traits>

#include "ne_internal.h"

namespace ne {

namespace {

constexpr int BX = 64;
constexpr int BY = 4;
// Rows of the deferring K1's blocks (the mixed form's FP32 iteration). Its columns are cheap, and
// blocks of 8 rows halve what each column pays of the block's fixed work (the column sort, the
// counters' reduction, the launch): 158 → 125 µs at 1440×560×10 against 4 rows, and 263 µs
// with 2 (same-box A/B; the FP64 K1 is 2 % slower with 8). A power of two.
#ifndef NE_FLUX_MIXED_ROWS
#define NE_FLUX_MIXED_ROWS 8
#endif

// Minimum wavefronts per SIMD requested for the flux-solve kernels (bounds their VGPRs).
#ifndef NE_FLUX_WAVES_PER_EU
#define NE_FLUX_WAVES_PER_EU 4
#endif

// 1: the mixed form (FP32 storage) starts the fixed point with the FP32 pre-iteration;
// 0: it runs the FP64 iteration of the FP64 form from the reference's initial guess.
#ifndef NE_MIXED_PREITERATION
#define NE_MIXED_PREITERATION 1
#endif

inline dim3 grid2d(int Nx, int Ny) { return dim3((Nx + BX - 1) / BX, (Ny + BY - 1) / BY); }

// Counting sort of a block's BX·BY columns by a small key (0 … SORT_BINS−1): returns the
// block-local column (0 … BX·BY−1) that work-item t computes, the columns taken in key order.
// Ties are ordered by the LDS atomics, which is immaterial (columns are independent).
constexpr int SORT_BINS = 64;
template <int N = BX * BY>
__device__ __forceinline__ int sorted_block_column(int key, int t) {
    __shared__ int s_hist[SORT_BINS];
    __shared__ int s_col[N];   // blocks of at most N work-items
    if (t < SORT_BINS) s_hist[t] = 0;
    __syncthreads();
    const int rank = atomicAdd(&s_hist[key], 1);
    __syncthreads();
    if (t < 64) {   // exclusive prefix over the bins: the first wavefront, a shuffle scan
        const int v = t < SORT_BINS ? s_hist[t] : 0;
        int x = v;
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (t >= d) x += y;
        }
        if (t < SORT_BINS) s_hist[t] = x - v;
    }
    __syncthreads();
    s_col[s_hist[key] + rank] = t;
    __syncthreads();
    return s_col[t];
}

// Sort key of a column for the flux solve: the iteration count its previous solve took, and
// the columns that ended stable (buoyancy flux into the ocean, ζ > 0: θ★ + 0.61·T·q★ > 0 with
// T ≈ 290 K) after the unstable ones, so that fewer wavefronts take both branches of the
// stability functions.
__device__ __forceinline__ int solve_order_key(const FluxOut& out, int i, int j) {
    const int it = min(max(out.iters.at2(i, j), 0), SORT_BINS / 2 - 2);
    const bool stable = out.ts.at2(i, j) + 176.0 * out.qs.at2(i, j) > 0.0;
    return stable ? SORT_BINS / 2 - 1 + it : it;
}

// Block-wide reduction of the solver counters, then one atomic per block (blocks of a
// power-of-two number of work-items, at most N).
template <int N = BX * BY>
__device__ void accumulate_counters(ne_counters* c, bool solved, int it, bool nonconv, bool nan) {
    __shared__ unsigned s_it[N], s_solved[N], s_nc[N], s_nan[N], s_max[N];
    int t = threadIdx.x + threadIdx.y * blockDim.x;
    int n = blockDim.x * blockDim.y;
    s_it[t] = solved ? (unsigned)it : 0u;
    s_solved[t] = solved ? 1u : 0u;
    s_nc[t] = (solved && nonconv) ? 1u : 0u;
    s_nan[t] = (solved && nan) ? 1u : 0u;
    s_max[t] = solved ? (unsigned)it : 0u;
    __syncthreads();
    for (int w = n / 2; w > 0; w >>= 1) {
        if (t < w) {
            s_it[t] += s_it[t + w];
            s_solved[t] += s_solved[t + w];
            s_nc[t] += s_nc[t + w];
            s_nan[t] += s_nan[t + w];
            s_max[t] = s_max[t] > s_max[t + w] ? s_max[t] : s_max[t + w];
        }
        __syncthreads();
    }
    if (t == 0 && c != nullptr && s_solved[0] > 0) {
        atomicAdd((unsigned long long*)&c->total_iterations, (unsigned long long)s_it[0]);
        atomicAdd(&c->solved_columns, s_solved[0]);
        if (s_nc[0]) atomicAdd(&c->nonconverged_columns, s_nc[0]);
        if (s_nan[0]) atomicAdd(&c->nan_columns, s_nan[0]);
        atomicMax(&c->max_iterations, s_max[0]);
    }
}

// The solver counters of a block, added once at the kernel's end from per-work-item counts (a
// persistent kernel sums them over its items first): each wavefront reduces its lanes
// (cross-lane), then one barrier, and the block's first work-item adds the wavefronts' partials
// to the global counters — one barrier where the tree over the block's work-items takes one per
// level. Blocks of N work-items, a multiple of 64; called by every work-item.
template <int N>
__device__ void add_block_counts(ne_counters* c, unsigned iters, unsigned solved, unsigned nonconv, unsigned nan,
                                 unsigned maxit) {
    constexpr int NW = N / 64;
    __shared__ unsigned s_part[5][NW];
    for (int d = 32; d >= 1; d >>= 1) {
        iters += (unsigned)__shfl_xor((int)iters, d, 64);
        solved += (unsigned)__shfl_xor((int)solved, d, 64);
        nonconv += (unsigned)__shfl_xor((int)nonconv, d, 64);
        nan += (unsigned)__shfl_xor((int)nan, d, 64);
        const unsigned o = (unsigned)__shfl_xor((int)maxit, d, 64);
        maxit = o > maxit ? o : maxit;
    }
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    const int w = t >> 6;
    if ((t & 63) == 0) {
        s_part[0][w] = iters;
        s_part[1][w] = solved;
        s_part[2][w] = nonconv;
        s_part[3][w] = nan;
        s_part[4][w] = maxit;
    }
    __syncthreads();
    if (t != 0 || c == nullptr) return;
    unsigned long long tot = 0;
    unsigned n = 0, nc = 0, nn = 0, m = 0;
    const int nw = ((int)(blockDim.x * blockDim.y) + 63) / 64;
    for (int k = 0; k < nw; k++) {
        tot += s_part[0][k];
        n += s_part[1][k];
        nc += s_part[2][k];
        nn += s_part[3][k];
        m = s_part[4][k] > m ? s_part[4][k] : m;
    }
    if (n == 0) return;
    atomicAdd((unsigned long long*)&c->total_iterations, tot);
    atomicAdd(&c->solved_columns, n);
    if (nc) atomicAdd(&c->nonconverged_columns, nc);
    if (nn) atomicAdd(&c->nan_columns, nn);
    atomicMax(&c->max_iterations, m);
}

// add_block_counts for a kernel with one column per work-item.
template <int N>
__device__ void accumulate_counters_once(ne_counters* c, bool solved, int it, bool nonconv, bool nan) {
    add_block_counts<N>(c, solved ? (unsigned)it : 0u, solved ? 1u : 0u, (solved && nonconv) ? 1u : 0u,
                        (solved && nan) ? 1u : 0u, solved ? (unsigned)it : 0u);
}

__device__ __forceinline__ void store_fluxes(const FluxOut& o, int i, int j, const ColumnFluxes& r) {
    o.Qv.at2(i, j) = r.Qv;
    o.Qt.at2(i, j) = r.Qt;
    o.Jv.at2(i, j) = r.Jv;
    o.rtx.at2(i, j) = r.rtx;
    o.rty.at2(i, j) = r.rty;
    o.us.at2(i, j) = r.us;
    o.ts.at2(i, j) = r.ts;
    o.qs.at2(i, j) = r.qs;
    o.Ts.at2(i, j) = r.Ts_C;
    if (o.iters.p) o.iters.at2(i, j) = r.iterations;
}

__device__ __forceinline__ bool finite3(double a, double b, double c) {
    return isfinite(a) && isfinite(b) && isfinite(c);
}

// ---------------------------------------------------------------------------

__global__ void immersed_levels_kernel(GridDev g, View3<const double> bottom, int32_t* kmin) {
    int i = blockIdx.x * BX + threadIdx.x;
    int j = blockIdx.y * BY + threadIdx.y;
    if (i >= g.Nx || j >= g.Ny) return;
    double h = bottom.at2(i, j);
    int k = 0;
    while (k < g.Nz && g.zc[k] <= h) k++;
    kmin[i + (int64_t)j * g.Nx] = k;
}

__global__ void fractional_indices_kernel(GridDev g, ne_source_grid s, double* fi, double* fj) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    double lc0 = s.lambda0 + 0.5 * s.dlambda;
    double pc0 = s.phi0 + 0.5 * s.dphi;
    if (n < g.Nx) fi[n] = s.topology_x == NE_FLAT ? 0.0 : (g.lambda_c[n] - lc0) / s.dlambda;
    if (n < g.Ny) fj[n] = s.topology_y == NE_FLAT ? 0.0 : (g.phi_c[n] - pc0) / s.dphi;
}

__global__ void log_kernel(const double* x, double* y, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = ne_log(x[i]);
}

__global__ void atan_ge1_kernel(const double* x, double* y, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = ne_atan_ge1(x[i]);
}

__global__ void exp_kernel(const double* x, double* y, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = ne_exp(x[i]);
}

__global__ void sqrt_kernel(const double* x, double* y, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = ne_sqrt(x[i]);
}

__global__ void cbrt_ge1_kernel(const double* x, double* y, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = ne_cbrt_ge1(x[i]);
}

__global__ void fractional_indices_nodes_kernel(GridDev g, const double* xn, int nx, int tx, const double* yn, int ny,
                                                int ty, double* fi, double* fj) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < g.Nx) fi[n] = fractional_index_nodes(g.lambda_c[n], xn, nx, tx);
    if (n < g.Ny) fj[n] = fractional_index_nodes(g.phi_c[n], yn, ny, ty);
}

// The ring of the per-phase surface entry points (grids with GridDev::ring): positions
// i = -1 … Nx on rows j = -1 and Ny, then i = -1 and Nx on rows 0 … Ny-1, in that order.
// The reference launches its interpolation and flux kernels over 0:N+1 in every non-Flat
// direction (InterfaceComputations.jl:99-116).
__host__ __device__ inline int ring_size(const GridDev& g) { return 2 * (g.Nx + 2) + 2 * g.Ny; }
__device__ __forceinline__ bool ring_position(const GridDev& g, int r, int& i, int& j) {
    const int W = g.Nx + 2;
    if (r < W) { i = r - 1; j = -1; return true; }
    r -= W;
    if (r < W) { i = r - 1; j = g.Ny; return true; }
    r -= W;
    if (r < g.Ny) { i = -1; j = r; return true; }
    r -= g.Ny;
    if (r < g.Ny) { i = g.Nx; j = r; return true; }
    return false;
}

// Fractional index of exchange node i along a direction of N nodes: the interior index, or for
// the ring nodes i = -1 and N, whose coordinates lie one spacing beyond the edge nodes of the
// equally spaced exchange grid, the index extended from the two edge nodes. In a periodic
// direction it is the wrapped node's index shifted by the source's period (the interpolation
// wraps it back); in a bounded one it lies beyond the source's edge node (the interpolation
// holds the edge value).
__device__ __forceinline__ double ring_index(const double* f, int i, int N) {
    if (i >= 0 && i < N) return f[i];
    if (N == 1) return f[0];
    return i < 0 ? f[0] - (f[1] - f[0]) : f[N - 1] + (f[N - 1] - f[N - 2]);
}

template <typename Real>
__device__ __forceinline__ void interp_atmos_at(const AtmosDev<Real>& a, const Neighbours& x, const Neighbours& y,
                                                int i, int j, View3<double>& u, View3<double>& v, View3<double>& T,
                                                View3<double>& p, View3<double>& q, View3<double>& rain,
                                                View3<double>& snow) {
    u.at2(i, j) = interp_field(a.u, x, y, a.time);
    v.at2(i, j) = interp_field(a.v, x, y, a.time);
    T.at2(i, j) = interp_field(a.T, x, y, a.time);
    p.at2(i, j) = interp_field(a.p, x, y, a.time);
    q.at2(i, j) = interp_field(a.q, x, y, a.time);
    if (rain.p) rain.at2(i, j) = interp_field(a.rain, x, y, a.time);
    if (snow.p) snow.at2(i, j) = interp_field(a.snow, x, y, a.time);
}

template <typename Real>
__global__ void interp_atmos_kernel(GridDev g, AtmosDev<Real> a, const double* fi, const double* fj,
                                    View3<double> u, View3<double> v, View3<double> T, View3<double> p,
                                    View3<double> q, View3<double> rain, View3<double> snow) {
    int i = blockIdx.x * BX + threadIdx.x;
    int j = blockIdx.y * BY + threadIdx.y;
    if (i >= g.Nx || j >= g.Ny) return;
    Neighbours x = neighbours(fi[i], a.Nx, a.tx);
    Neighbours y = neighbours(fj[j], a.Ny, a.ty);
    interp_atmos_at(a, x, y, i, j, u, v, T, p, q, rain, snow);
}

template <typename Real>
__global__ void interp_atmos_ring_kernel(GridDev g, AtmosDev<Real> a, const double* fi, const double* fj,
                                         View3<double> u, View3<double> v, View3<double> T, View3<double> p,
                                         View3<double> q, View3<double> rain, View3<double> snow) {
    int i, j;
    if (!ring_position(g, (int)(blockIdx.x * blockDim.x + threadIdx.x), i, j)) return;
    Neighbours x = neighbours(ring_index(fi, i, g.Nx), a.Nx, a.tx);
    Neighbours y = neighbours(ring_index(fj, j, g.Ny), a.Ny, a.ty);
    interp_atmos_at(a, x, y, i, j, u, v, T, p, q, rain, snow);
}

template <typename Real>
__global__ void interp_radiation_kernel(GridDev g, RadDev<Real> r, const double* fi, const double* fj,
                                        View3<double> sw, View3<double> lw) {
    int i = blockIdx.x * BX + threadIdx.x;
    int j = blockIdx.y * BY + threadIdx.y;
    if (i >= g.Nx || j >= g.Ny) return;
    Neighbours x = neighbours(fi[i], r.Nx, r.tx);
    Neighbours y = neighbours(fj[j], r.Ny, r.ty);
    sw.at2(i, j) = interp_field(r.sw, x, y, r.time);
    lw.at2(i, j) = interp_field(r.lw, x, y, r.time);
}

template <typename Real>
__global__ void interp_radiation_ring_kernel(GridDev g, RadDev<Real> r, const double* fi, const double* fj,
                                             View3<double> sw, View3<double> lw) {
    int i, j;
    if (!ring_position(g, (int)(blockIdx.x * blockDim.x + threadIdx.x), i, j)) return;
    Neighbours x = neighbours(ring_index(fi, i, g.Nx), r.Nx, r.tx);
    Neighbours y = neighbours(ring_index(fj, j, g.Ny), r.Ny, r.ty);
    sw.at2(i, j) = interp_field(r.sw, x, y, r.time);
    lw.at2(i, j) = interp_field(r.lw, x, y, r.time);
}

// Ocean surface values of a column: T, S at k = Nz and ℑxᶜᵃᵃ u, ℑyᵃᶜᵃ v
// (atmosphere_ocean_fluxes.jl:62-71).
template <typename Real>
__device__ __forceinline__ void ocean_surface(const GridDev& g, const View3<const Real>& T, const View3<const Real>& S,
                                              const View3<const Real>& u, const View3<const Real>& v, int i, int j,
                                              double& To, double& So, double& uo, double& vo) {
    int kN = g.Nz - 1;
    To = (double)T(i, j, kN);
    So = (double)S(i, j, kN);
    int ip = g.tx == NE_PERIODIC ? wrap_near(i + 1, g.Nx) : i + 1;
    int jp = g.ty == NE_PERIODIC ? wrap_near(j + 1, g.Ny) : j + 1;
    uo = ((double)u(i, j, kN) + (double)u(ip, j, kN)) / 2.0;
    vo = ((double)v(i, j, kN) + (double)v(i, jp, kN)) / 2.0;
}

template <bool Fast, bool Consts>
__global__ void __launch_bounds__(BX* BY) __attribute__((amdgpu_waves_per_eu(NE_FLUX_WAVES_PER_EU)))
flux_solve_exchange_kernel(GridDev g, ne_params p, View3<const double> ua, View3<const double> va,
                                           View3<const double> Ta, View3<const double> pa, View3<const double> qa,
                                           View3<const double> T, View3<const double> S, View3<const double> u,
                                           View3<const double> v, View3<const double> sw,
                                           View3<const double> lw, FluxOut out, ne_counters* counters) {
    int i = blockIdx.x * BX + threadIdx.x;
    int j = blockIdx.y * BY + threadIdx.y;
    if (out.iters.p) {   // columns in the order K1 takes them (previous iteration count, stability)
        const bool own = i < g.Nx && j < g.Ny;
        const int key = own ? solve_order_key(out, i, j) : SORT_BINS - 1;
        const int c = sorted_block_column(key, threadIdx.x + threadIdx.y * BX);
        i = blockIdx.x * BX + c % BX;
        j = blockIdx.y * BY + c / BX;
    }
    bool in = i < g.Nx && j < g.Ny;
    ColumnFluxes r;
    bool inactive = true;
    if (in) {
        double To, So, uo, vo;
        ocean_surface<double>(g, T, S, u, v, i, j, To, So, uo, vo);
        inactive = !g.active(i, j, g.Nz - 1);
        double SW = 0.0, LW = 0.0;
        if (!Fast && p.interface_temperature == NE_TEMPERATURE_SKIN && p.radiation) {
            SW = sw.at2(i, j);
            LW = lw.at2(i, j);
        }
        r = flux_column<Fast, false, Consts>(p, ua.at2(i, j), va.at2(i, j), Ta.at2(i, j), pa.at2(i, j), qa.at2(i, j), To, uo, vo,
                              inactive, SW, LW);
        store_fluxes(out, i, j, r);
    }
    bool solved = in && !inactive;
    accumulate_counters_once<BX * BY>(counters, solved, solved ? r.iterations : 0, solved && !r.converged,
                                      solved && !finite3(r.us, r.ts, r.qs));
}

// The flux solve on the ring: the exchange state is read at the ring position (the
// interpolation's ring), the ocean surface of a ring node in a periodic direction is that of
// the wrapped interior column (the values the reference's filled halo holds), and a ring node
// beyond a bounded edge is inactive (zero interface state). The solve runs as in the interior
// (same instance of the column solve); the ring is not counted in the solver counters.
template <bool Fast, bool Consts>
__global__ void __launch_bounds__(BX* BY) __attribute__((amdgpu_waves_per_eu(NE_FLUX_WAVES_PER_EU)))
flux_solve_ring_kernel(GridDev g, ne_params p, View3<const double> ua, View3<const double> va,
                       View3<const double> Ta, View3<const double> pa, View3<const double> qa,
                       View3<const double> T, View3<const double> S, View3<const double> u,
                       View3<const double> v, View3<const double> sw, View3<const double> lw, FluxOut out) {
    int i, j;
    if (!ring_position(g, (int)(blockIdx.x * blockDim.x + threadIdx.x), i, j)) return;
    double To = 0.0, So = 0.0, uo = 0.0, vo = 0.0;
    bool inactive = true;
    if (g.column(i, j) >= 0) {
        const int wi = g.wx(i), wj = g.wy(j);
        ocean_surface<double>(g, T, S, u, v, wi, wj, To, So, uo, vo);
        inactive = !g.active(wi, wj, g.Nz - 1);
    }
    double SW = 0.0, LW = 0.0;
    if (!Fast && p.interface_temperature == NE_TEMPERATURE_SKIN && p.radiation) {
        SW = sw.at2(i, j);
        LW = lw.at2(i, j);
    }
    ColumnFluxes r = flux_column<Fast, false, Consts>(p, ua.at2(i, j), va.at2(i, j), Ta.at2(i, j), pa.at2(i, j),
                                                      qa.at2(i, j), To, uo, vo, inactive, SW, LW);
    store_fluxes(out, i, j, r);
}

// One column of K1: interpolate the atmosphere at the column, solve, store the fluxes.
// The skin-temperature balance also reads the downwelling radiation, interpolated here
// exactly as K2 interpolates it.
template <typename Real, bool Fast, bool Consts, bool Defer = false>
__device__ __forceinline__ ColumnFluxes surface_solve_column(const GridDev& g, const ne_params& p,
                                                             const AtmosDev<Real>& a, const double* __restrict__ fi,
                                                             const double* __restrict__ fj, const RadDev<Real>& rd,
                                                             const double* __restrict__ rfi,
                                                             const double* __restrict__ rfj,
                                                             const View3<const Real>& T, const View3<const Real>& S,
                                                             const View3<const Real>& u, const View3<const Real>& v,
                                                             const FluxOut& out, int i, int j, bool& inactive) {
    Neighbours x = neighbours(fi[i], a.Nx, a.tx);
    Neighbours y = neighbours(fj[j], a.Ny, a.ty);
    double ua = interp_field(a.u, x, y, a.time);
    double va = interp_field(a.v, x, y, a.time);
    double Ta = interp_field(a.T, x, y, a.time);
    double pa = interp_field(a.p, x, y, a.time);
    double qa = interp_field(a.q, x, y, a.time);
    double To, So, uo, vo;
    ocean_surface<Real>(g, T, S, u, v, i, j, To, So, uo, vo);
    inactive = !g.active(i, j, g.Nz - 1);
    double SW = 0.0, LW = 0.0;
    if (!Fast && p.interface_temperature == NE_TEMPERATURE_SKIN && p.radiation) {
        Neighbours rx = neighbours(rfi[i], rd.Nx, rd.tx);
        Neighbours ry = neighbours(rfj[j], rd.Ny, rd.ty);
        SW = interp_field(rd.sw, rx, ry, rd.time);
        LW = interp_field(rd.lw, rx, ry, rd.time);
    }
    // FP32 storage (the mixed form) also takes the FP32 pre-iteration of the fixed point
    ColumnFluxes r = flux_column<Fast, NE_MIXED_PREITERATION && sizeof(Real) == 4, Consts, Defer>(
        p, ua, va, Ta, pa, qa, To, uo, vo, inactive, SW, LW);
    if (!Defer || !r.deferred) store_fluxes(out, i, j, r);
    return r;
}


// Freeze clamp of one column (freezing_limited_ocean_temperature.jl:95-118); returns frazil heat.
template <typename Real>
__device__ __forceinline__ double freeze_column(const GridDev& g, const ne_params& p, const View3<Real>& T,
                                                const View3<const Real>& S, int i, int j, double dt) {
    double rho = p.ocean_reference_density, c = p.ocean_heat_capacity;
    double dQ = 0.0;
    for (int k = g.Nz - 1; k >= 0; k--) {
        double dz = g.dz[k];
        double Tk = (double)T(i, j, k), Sk = (double)S(i, j, k);
        double Tm = p.freshwater_melting_temperature - p.liquidus_slope * Sk;
        bool freezing = Tk < Tm;
        double dE = freezing ? rho * c * (Tm - Tk) : 0.0;
        if (freezing) T(i, j, k) = (Real)Tm;
        dQ = dQ - dE * dz / dt;
    }
    return dQ;
}

__global__ void freeze_clamp_kernel(GridDev g, ne_params p, View3<double> T, View3<const double> S,
                                    View3<double> frazil, double dt) {
    int i = blockIdx.x * BX + threadIdx.x;
    int j = blockIdx.y * BY + threadIdx.y;
    if (i >= g.Nx || j >= g.Ny) return;
    frazil.at2(i, j) = freeze_column<double>(g, p, T, S, i, j, dt);
}

// Net ocean fluxes of one column (assemble_net_ocean_fluxes.jl:74-153, ℵ = 0).
struct NetColumn {
    double tau_x, tau_y, J_T, J_S, J_w, J_H;
};

// τ on the face between a cell and its west (south) neighbour: the face average of
// (1 − ℵ)ρτ/ρₒ of the two cells plus the sea-ice term, 0 here (assemble_net_ocean_fluxes.jl:139-142).
__device__ __forceinline__ double tau_face(const ne_params& p, double r, double r_nb) {
    double rinv = 1.0 / p.ocean_reference_density;
    double aice = 0.0;
    double t_ao = (rinv * (1.0 - aice) * r + rinv * (1.0 - aice) * r_nb) / 2.0;
    double t_io = 0.0 * rinv * 0.0;
    return t_ao + t_io;
}

__device__ __forceinline__ NetColumn assemble_column(const GridDev& g, const ne_params& p, double Qt, double Qv,
                                                     double Jv, double rtx, double rtx_w, double rty,
                                                     double rty_s, double Jr, double Js, double To,
                                                     bool inactive) {
    double rinv = 1.0 / p.ocean_reference_density;
    double cinv = 1.0 / p.ocean_heat_capacity;
    double aice = 0.0, Jl = 0.0, Ps = 0.0;
    double SQao = (Qt + Qv) * (1.0 - aice);
    double SFao = -(Jr + Jl + Js - Ps) * rinv + (1.0 - aice) * Jv * rinv;
    double Jwao = -SFao;
    double Qin = 0.0, Jsio = 0.0, Jwio = 0.0;
    double JTao = SQao * rinv * cinv;
    double JTio = Qin * rinv * cinv;
    NetColumn n;
    n.tau_x = inactive ? 0.0 : tau_face(p, rtx, rtx_w);
    n.tau_y = inactive ? 0.0 : tau_face(p, rty, rty_s);
    n.J_T = inactive ? 0.0 : JTao + JTio;
    n.J_S = inactive ? 0.0 : Jsio;
    n.J_w = inactive ? 0.0 : Jwao + Jwio;
    n.J_H = inactive ? 0.0 : To * Jwao;
    return n;
}

// Radiative part (apply_air_sea_radiative_fluxes.jl:62-111): returns Jᵀ_rad, writes diagnostics.
__device__ __forceinline__ double radiation_column(const ne_params& p, double Ts_C, double SW, double LW,
                                                   double& Iup, double& Ila, double& Itsw) {
    double rinv = 1.0 / p.ocean_reference_density;
    double cinv = 1.0 / p.ocean_heat_capacity;
    double aice = 0.0;
    double Ts = Ts_C + NE_CELSIUS_TO_KELVIN;
    // T^4 as Julia evaluates Float64^Integer (radiation_kernels.jl:3): the compensated power, as in
    // the skin-temperature balance
    Iup = p.stefan_boltzmann_constant * p.ocean_emissivity * pow_int(Ts, 4);
    Ila = -p.ocean_emissivity * LW;
    Itsw = -(1.0 - p.ocean_albedo) * SW;
    Ila = Ila * (1.0 - aice);
    Itsw = Itsw * (1.0 - aice);
    double Iup_ocean = Iup * (1.0 - aice);
    // shortwave_radiative_forcing (radiative_forcing.jl:84-91): with TwoColorRadiation the
    // transmitted shortwave becomes its surface flux J₀ and leaves Jᵀ
    double Qss = p.radiative_forcing == NE_RADIATIVE_FORCING_TWO_COLOR ? 0.0 : Itsw;
    double SQ = Iup_ocean + Ila + Qss;
    return SQ * rinv * cinv;
}

__device__ __forceinline__ void store_penetrating_shortwave(const ne_params& p, const NetOut& n, int i, int j,
                                                            double Itsw) {
    if (p.radiative_forcing == NE_RADIATIVE_FORCING_TWO_COLOR)
        n.J_sw.at2(i, j) = -Itsw / (p.ocean_reference_density * p.ocean_heat_capacity);
}

// Momentum fluxes of the west / south neighbours (zero state outside the domain).
__device__ __forceinline__ void neighbour_momentum(const GridDev& g, const View3<double>& rtx,
                                                   const View3<double>& rty, int i, int j, double& rtx_w,
                                                   double& rty_s) {
    rtx_w = g.column(i - 1, j) >= 0 ? rtx.at2(g.wx(i - 1), j) : 0.0;
    rty_s = g.column(i, j - 1) >= 0 ? rty.at2(i, g.wy(j - 1)) : 0.0;
}

__global__ void assemble_kernel(GridDev g, ne_params p, View3<double> Qv, View3<double> Qt, View3<double> Jv,
                                View3<double> rtx, View3<double> rty, View3<double> rain, View3<double> snow,
                                View3<const double> sst, NetOut n) {
    int i = blockIdx.x * BX + threadIdx.x;
    int j = blockIdx.y * BY + threadIdx.y;
    if (i >= g.Nx || j >= g.Ny) return;
    double rtx_w, rty_s;
    neighbour_momentum(g, rtx, rty, i, j, rtx_w, rty_s);
    double Jr = rain.p ? rain.at2(i, j) : 0.0;
    double Js = snow.p ? snow.at2(i, j) : 0.0;
    bool inactive = !g.active(i, j, g.Nz - 1);
    NetColumn c = assemble_column(g, p, Qt.at2(i, j), Qv.at2(i, j), Jv.at2(i, j), rtx.at2(i, j), rtx_w,
                                  rty.at2(i, j), rty_s, Jr, Js, sst(i, j, g.Nz - 1), inactive);
    n.tau_x.at2(i, j) = c.tau_x;
    n.tau_y.at2(i, j) = c.tau_y;
    n.J_T.at2(i, j) = c.J_T;
    n.J_S.at2(i, j) = c.J_S;
    n.J_w.at2(i, j) = c.J_w;
    n.J_H.at2(i, j) = c.J_H;
}

__global__ void radiation_kernel(GridDev g, ne_params p, View3<const double> sw, View3<const double> lw,
                                 View3<const double> Ts, NetOut n, RadOut r) {
    int i = blockIdx.x * BX + threadIdx.x;
    int j = blockIdx.y * BY + threadIdx.y;
    if (i >= g.Nx || j >= g.Ny) return;
    double Iup, Ila, Itsw;
    double JTrad = radiation_column(p, Ts.at2(i, j), sw.at2(i, j), lw.at2(i, j), Iup, Ila, Itsw);
    bool inactive = !g.active(i, j, g.Nz - 1);
    n.J_T.at2(i, j) = n.J_T.at2(i, j) + (inactive ? 0.0 : JTrad);
    store_penetrating_shortwave(p, n, i, j, Itsw);
    if (r.up_lw.p) r.up_lw.at2(i, j) = Iup;
    if (r.down_lw.p) r.down_lw.at2(i, j) = -Ila;
    if (r.down_sw.p) r.down_sw.at2(i, j) = -Itsw;
}

// One column of K2: freeze clamp, net-flux assembly and radiation.
template <typename Real, bool TAU>
__device__ __forceinline__ void surface_finalize_column(const GridDev& g, const ne_params& p, const AtmosDev<Real>& a,
                                                        const double* __restrict__ afi,
                                                        const double* __restrict__ afj, const RadDev<Real>& rd,
                                                        const double* __restrict__ rfi,
                                                        const double* __restrict__ rfj, const View3<Real>& T,
                                                        const View3<const Real>& S, const FluxOut& f,
                                                        const NetOut& n, const RadOut& ro,
                                                        const View3<double>& frazil, double dt_frazil,
                                                        const uint8_t* __restrict__ freeze_flags, double Qt,
                                                        double Qv, double Jv, double Ts_C, int i, int j) {
    // after a step, only columns its last stage flagged can hold a cell below freezing; the
    // others would run the clamp without a change and with dQ = 0 (the same +0.0)
    const bool visit = freeze_flags == nullptr || freeze_flags[(int64_t)i + (int64_t)j * g.Nx] != 0;
    double dQ = visit ? freeze_column<Real>(g, p, T, S, i, j, dt_frazil) : 0.0;
    if (frazil.p) frazil.at2(i, j) = dQ;

    Neighbours x = neighbours(afi[i], a.Nx, a.tx);
    Neighbours y = neighbours(afj[j], a.Ny, a.ty);
    double Jr = interp_field(a.rain, x, y, a.time);
    double Js = interp_field(a.snow, x, y, a.time);
    double rtx = 0.0, rty = 0.0, rtx_w = 0.0, rty_s = 0.0;   // τ here, or in the τ kernel after all columns
    if (TAU) {
        rtx = f.rtx.at2(i, j);
        rty = f.rty.at2(i, j);
        neighbour_momentum(g, f.rtx, f.rty, i, j, rtx_w, rty_s);
    }
    bool inactive = !g.active(i, j, g.Nz - 1);
    NetColumn c = assemble_column(g, p, Qt, Qv, Jv, rtx, rtx_w, rty, rty_s, Jr, Js, (double)T(i, j, g.Nz - 1),
                                  inactive);
    double JT = c.J_T;
    if (p.radiation) {
        Neighbours rx = neighbours(rfi[i], rd.Nx, rd.tx);
        Neighbours ry = neighbours(rfj[j], rd.Ny, rd.ty);
        double SW = interp_field(rd.sw, rx, ry, rd.time);
        double LW = interp_field(rd.lw, rx, ry, rd.time);
        double Iup, Ila, Itsw;
        double JTrad = radiation_column(p, Ts_C, SW, LW, Iup, Ila, Itsw);
        JT = JT + (inactive ? 0.0 : JTrad);
        store_penetrating_shortwave(p, n, i, j, Itsw);
        if (ro.up_lw.p) ro.up_lw.at2(i, j) = Iup;
        if (ro.down_lw.p) ro.down_lw.at2(i, j) = -Ila;
        if (ro.down_sw.p) ro.down_sw.at2(i, j) = -Itsw;
    }
    if (TAU) {
        n.tau_x.at2(i, j) = c.tau_x;
        n.tau_y.at2(i, j) = c.tau_y;
    }
    n.J_T.at2(i, j) = JT;
    n.J_S.at2(i, j) = c.J_S;
    n.J_w.at2(i, j) = c.J_w;
    n.J_H.at2(i, j) = c.J_H;
}

// K1 of the fused update_state: per column, interpolate the atmosphere and solve the fluxes,
// then (with the solved values still in registers) the freeze clamp, the net fluxes other
// than τ and the radiation, in the reference's phase order (fluxes from the unclamped SST,
// clamp, assembly with the clamped SST, radiation).
// Defer (FP32 storage with parameters that take the FP32 iteration): a column whose FP32
// iteration does not finish it is appended to `deferred` (count, then flattened columns) and left
// to surface_deferred_kernel.
template <typename Real, bool Fast, bool Consts, bool Defer = false, int KBY = BY>
__global__ void __launch_bounds__(BX* KBY) __attribute__((amdgpu_waves_per_eu(NE_FLUX_WAVES_PER_EU)))
surface_update_kernel(GridDev g, ne_params p, AtmosDev<Real> a, const double* __restrict__ fi,
                      const double* __restrict__ fj, RadDev<Real> rd, const double* __restrict__ rfi,
                      const double* __restrict__ rfj, View3<Real> T, View3<const Real> S, View3<const Real> u,
                      View3<const Real> v, FluxOut out, NetOut n, RadOut ro, View3<double> frazil, double dt_frazil,
                      const uint8_t* __restrict__ freeze_flags, ne_counters* counters, int* deferred) {
    int i = blockIdx.x * BX + threadIdx.x;
    int j = blockIdx.y * KBY + threadIdx.y;
    if (out.iters.p && !Defer) {
        // (Not for the mixed form's FP32 iteration, whose counts vary less: without the sort and
        // its barriers its K1 took 124 → 117 µs at 1440×560×10.)
        // Columns of the block in order of the iteration count the previous solve took, the
        // stable ones last (solve_order_key): the fixed point runs until the slowest column of a
        // wavefront converges, so wavefronts of columns with like counts leave fewer lanes idle,
        // and a wavefront without stable columns skips that branch. Each column's result does
        // not depend on the lane that computes it.
        const int t = threadIdx.x + threadIdx.y * BX;
        const bool own = i < g.Nx && j < g.Ny;
        const int key = own ? solve_order_key(out, i, j) : SORT_BINS - 1;
        const int c = sorted_block_column<BX * KBY>(key, t);
        i = blockIdx.x * BX + c % BX;
        j = blockIdx.y * KBY + c / BX;
    }
    bool in = i < g.Nx && j < g.Ny;
    ColumnFluxes r;
    r.deferred = false;
    bool inactive = true;
    if (in) {
        r = surface_solve_column<Real, Fast, Consts, Defer>(g, p, a, fi, fj, rd, rfi, rfj, cview(T), S, u, v, out, i,
                                                            j, inactive);
        if (Defer && r.deferred)
            deferred[1 + atomicAdd(&deferred[0], 1)] = i + j * g.Nx;
        else
            surface_finalize_column<Real, false>(g, p, a, fi, fj, rd, rfi, rfj, T, S, out, n, ro, frazil, dt_frazil,
                                                 freeze_flags, r.Qt, r.Qv, r.Jv, r.Ts_C, i, j);
    }
    bool solved = in && !inactive && !(Defer && r.deferred);
    accumulate_counters_once<BX * KBY>(counters, solved, solved ? r.iterations : 0, solved && !r.converged,
                        solved && !finite3(r.us, r.ts, r.qs));
}

// The columns surface_update_kernel deferred, finished as the kernel without deferral finishes
// them (the FP32 iteration again, then the FP64 one): a grid-stride loop over the list, which is
// empty in most steps. surface_tau_kernel, which runs after it, clears the count.
template <typename Real, bool Fast, bool Consts>
__global__ void __launch_bounds__(BX* BY) __attribute__((amdgpu_waves_per_eu(NE_FLUX_WAVES_PER_EU)))
surface_deferred_kernel(GridDev g, ne_params p, AtmosDev<Real> a, const double* __restrict__ fi,
                        const double* __restrict__ fj, RadDev<Real> rd, const double* __restrict__ rfi,
                        const double* __restrict__ rfj, View3<Real> T, View3<const Real> S, View3<const Real> u,
                        View3<const Real> v, FluxOut out, NetOut n, RadOut ro, View3<double> frazil,
                        double dt_frazil, const uint8_t* __restrict__ freeze_flags, ne_counters* counters,
                        const int* __restrict__ deferred) {
    const int count = deferred[0];   // the same for every work-item: nothing writes it meanwhile
    // rounds of one column per work-item, uniform over the block (the counters synchronise it)
    for (int base = blockIdx.x * BX * BY; base < count; base += gridDim.x * BX * BY) {
        const int q = base + (int)(threadIdx.x + threadIdx.y * BX);
        ColumnFluxes r;
        bool inactive = true;
        if (q < count) {
            const int col = deferred[1 + q];
            const int i = col % g.Nx, j = col / g.Nx;
            r = surface_solve_column<Real, Fast, Consts>(g, p, a, fi, fj, rd, rfi, rfj, cview(T), S, u, v, out, i, j,
                                                         inactive);
            surface_finalize_column<Real, false>(g, p, a, fi, fj, rd, rfi, rfj, T, S, out, n, ro, frazil,
                                                 dt_frazil, freeze_flags, r.Qt, r.Qv, r.Jv, r.Ts_C, i, j);
        }
        const bool solved = q < count && !inactive;
        accumulate_counters(counters, solved, solved ? r.iterations : 0, solved && !r.converged,
                            solved && !finite3(r.us, r.ts, r.qs));
    }
}

// After the fused K1: τ of every column from its own and its west / south neighbours'
// momentum fluxes, which K1 wrote (the one part of the update that reads other columns).
__global__ void __launch_bounds__(BX* BY) surface_tau_kernel(GridDev g, ne_params p, FluxOut f, NetOut n,
                                                            int* deferred) {
    int i = blockIdx.x * BX + threadIdx.x;
    int j = blockIdx.y * BY + threadIdx.y;
    if (deferred && i == 0 && j == 0) deferred[0] = 0;   // the deferred columns have been finished
    if (i >= g.Nx || j >= g.Ny) return;
    double rtx_w, rty_s;
    neighbour_momentum(g, f.rtx, f.rty, i, j, rtx_w, rty_s);
    bool inactive = !g.active(i, j, g.Nz - 1);
    n.tau_x.at2(i, j) = inactive ? 0.0 : tau_face(p, f.rtx.at2(i, j), rtx_w);
    n.tau_y.at2(i, j) = inactive ? 0.0 : tau_face(p, f.rty.at2(i, j), rty_s);
}

// Ensemble form of K1: a persistent grid walks the (member, first column) work list; an
// item is blockDim.x consecutive columns (i fastest) of one member. Members with the
// default Edson constants are listed for the shared-ψ instance (Fast), the others for
// the general one, as the single-model launcher chooses.
// FUSED (the step): each column then goes on as in the single model's K1 (freeze clamp of the
// flagged columns, net fluxes other than τ, radiation); without it (the autotuner's timing
// runs) the kernel only solves, leaving the state as it is.
template <typename Real, bool Fast, bool Consts, bool FUSED>
__global__ void __launch_bounds__(BX* BY) __attribute__((amdgpu_waves_per_eu(NE_FLUX_WAVES_PER_EU)))
batch_surface_solve_kernel(const BatchRec<Real>* __restrict__ recs, const WorkItem* __restrict__ work, int nwork,
                           double dt_frazil, ne_counters* counters) {
    // this work-item's counts over its items, added to the counters once at the end
    unsigned n_it = 0u, n_solved = 0u, n_nc = 0u, n_nan = 0u, n_max = 0u;
    for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
        const WorkItem it = work[wi];
        const BatchRec<Real>& rc = recs[it.member];
        const GridDev g = rc.g;
        int col = it.index + (int)threadIdx.x;
        if (rc.f.iters.p && blockDim.x > 64 && !(sizeof(Real) == 4 && Fast)) {
            // the item's columns in the single model's order (previous iteration count, stability;
            // the mixed form's FP32 iteration, as there, takes them unsorted)
            const int key = col < g.Nx * g.Ny ? solve_order_key(rc.f, col % g.Nx, col / g.Nx) : SORT_BINS - 1;
            col = it.index + sorted_block_column(key, (int)threadIdx.x);
        }
        const bool in = col < g.Nx * g.Ny;
        const int i = in ? col % g.Nx : 0, j = in ? col / g.Nx : 0;
        ColumnFluxes r;
        bool inactive = true;
        if (in)
        {
            r = surface_solve_column<Real, Fast, Consts>(g, rc.p, rc.a, rc.afi, rc.afj, rc.r, rc.rfi, rc.rfj, cview(rc.T),
                                                 cview(rc.S), rc.u, rc.v, rc.f, i, j, inactive);
            if (FUSED)
                surface_finalize_column<Real, false>(g, rc.p, rc.a, rc.afi, rc.afj, rc.r, rc.rfi, rc.rfj, rc.T,
                                                     cview(rc.S), rc.f, rc.n, rc.ro, rc.frazil, dt_frazil,
                                                     rc.freeze_flags, r.Qt, r.Qv, r.Jv, r.Ts_C, i, j);
        }
        if (in && !inactive) {
            n_it += (unsigned)r.iterations;
            n_solved += 1u;
            n_nc += r.converged ? 0u : 1u;
            n_nan += finite3(r.us, r.ts, r.qs) ? 0u : 1u;
            n_max = (unsigned)r.iterations > n_max ? (unsigned)r.iterations : n_max;
        }
    }
    add_block_counts<BX * BY>(counters, n_it, n_solved, n_nc, n_nan, n_max);
}

// Ensemble form of K2 over the same work list as the ensemble K1: τ of every column.
template <typename Real>
__global__ void __launch_bounds__(BX* BY)
batch_surface_tau_kernel(const BatchRec<Real>* __restrict__ recs, const WorkItem* __restrict__ work, int nwork) {
    for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
        const WorkItem it = work[wi];
        const BatchRec<Real>& rc = recs[it.member];
        const GridDev g = rc.g;
        const int col = it.index + (int)threadIdx.x;
        if (col >= g.Nx * g.Ny) continue;
        const int i = col % g.Nx, j = col / g.Nx;
        double rtx_w, rty_s;
        neighbour_momentum(g, rc.f.rtx, rc.f.rty, i, j, rtx_w, rty_s);
        const bool inactive = !g.active(i, j, g.Nz - 1);
        rc.n.tau_x.at2(i, j) = inactive ? 0.0 : tau_face(rc.p, rc.f.rtx.at2(i, j), rtx_w);
        rc.n.tau_y.at2(i, j) = inactive ? 0.0 : tau_face(rc.p, rc.f.rty.at2(i, j), rty_s);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

template <typename Real>
AtmosDev<Real> atmos_dev(const ne_prescribed_atmosphere& a) {
    AtmosDev<Real> d;
    d.Nx = a.grid.Nx; d.Ny = a.grid.Ny; d.tx = a.grid.topology_x; d.ty = a.grid.topology_y;
    d.u = view<const Real>(a.u); d.v = view<const Real>(a.v); d.T = view<const Real>(a.T);
    d.q = view<const Real>(a.q); d.p = view<const Real>(a.p);
    d.rain = view<const Real>(a.rain); d.snow = view<const Real>(a.snow);
    d.time = a.time;
    return d;
}

template <typename Real>
RadDev<Real> rad_dev(const ne_prescribed_radiation& r) {
    RadDev<Real> d;
    d.Nx = r.grid.Nx; d.Ny = r.grid.Ny; d.tx = r.grid.topology_x; d.ty = r.grid.topology_y;
    d.sw = view<const Real>(r.downwelling_shortwave);
    d.lw = view<const Real>(r.downwelling_longwave);
    d.time = r.time;
    return d;
}

template AtmosDev<double> atmos_dev<double>(const ne_prescribed_atmosphere&);
template AtmosDev<float> atmos_dev<float>(const ne_prescribed_atmosphere&);
template RadDev<double> rad_dev<double>(const ne_prescribed_radiation&);
template RadDev<float> rad_dev<float>(const ne_prescribed_radiation&);

FluxOut flux_out(const ne_interface_fluxes& f) {
    FluxOut o;
    o.Qv = view<double>(f.latent_heat); o.Qt = view<double>(f.sensible_heat); o.Jv = view<double>(f.water_vapor);
    o.rtx = view<double>(f.x_momentum); o.rty = view<double>(f.y_momentum);
    o.us = view<double>(f.friction_velocity); o.ts = view<double>(f.temperature_scale);
    o.qs = view<double>(f.water_vapor_scale); o.Ts = view<double>(f.interface_temperature);
    o.iters = view<int32_t>(f.iterations);
    return o;
}

NetOut net_out(const ne_net_ocean_fluxes& n) {
    NetOut o;
    o.tau_x = view<double>(n.tau_x); o.tau_y = view<double>(n.tau_y); o.J_T = view<double>(n.J_T);
    o.J_S = view<double>(n.J_S); o.J_w = view<double>(n.J_w); o.J_H = view<double>(n.J_H);
    o.J_sw = view<double>(n.shortwave_surface_flux);
    return o;
}

RadOut rad_out(const ne_radiative_fluxes& r) {
    RadOut o;
    o.up_lw = view<double>(r.upwelling_longwave);
    o.down_lw = view<double>(r.downwelling_longwave);
    o.down_sw = view<double>(r.downwelling_shortwave);
    return o;
}

hipError_t launch_immersed_levels(const GridDev& g, View3<const double> bottom, int32_t* kmin, hipStream_t s) {
    hipLaunchKernelGGL(immersed_levels_kernel, grid2d(g.Nx, g.Ny), dim3(BX, BY), 0, s, g, bottom, kmin);
    return hipGetLastError();
}

hipError_t launch_exp(const double* x, double* y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(exp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t launch_sqrt(const double* x, double* y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(sqrt_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t launch_cbrt_ge1(const double* x, double* y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(cbrt_ge1_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t launch_atan_ge1(const double* x, double* y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(atan_ge1_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t launch_log(const double* x, double* y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(log_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t launch_fractional_indices_nodes(const GridDev& g, const double* xn, int nx, int tx, const double* yn,
                                          int ny, int ty, double* fi, double* fj, hipStream_t s) {
    int n = g.Nx > g.Ny ? g.Nx : g.Ny;
    hipLaunchKernelGGL(fractional_indices_nodes_kernel, dim3((n + 255) / 256), dim3(256), 0, s, g, xn, nx, tx, yn, ny,
                       ty, fi, fj);
    return hipGetLastError();
}

hipError_t launch_fractional_indices(const GridDev& g, const ne_source_grid& src, double* fi, double* fj,
                                     hipStream_t s) {
    int n = g.Nx > g.Ny ? g.Nx : g.Ny;
    hipLaunchKernelGGL(fractional_indices_kernel, dim3((n + 255) / 256), dim3(256), 0, s, g, src, fi, fj);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_interp_atmos(const GridDev& g, const AtmosDev<Real>& a, const double* fi, const double* fj,
                               const ne_exchange_state& o, hipStream_t s) {
    hipLaunchKernelGGL(interp_atmos_kernel<Real>, grid2d(g.Nx, g.Ny), dim3(BX, BY), 0, s, g, a, fi, fj,
                       view<double>(o.u), view<double>(o.v), view<double>(o.T), view<double>(o.p),
                       view<double>(o.q), view<double>(o.rain), view<double>(o.snow));
    if (g.ring)
        hipLaunchKernelGGL(interp_atmos_ring_kernel<Real>, dim3((ring_size(g) + 255) / 256), dim3(256), 0, s, g, a,
                           fi, fj, view<double>(o.u), view<double>(o.v), view<double>(o.T), view<double>(o.p),
                           view<double>(o.q), view<double>(o.rain), view<double>(o.snow));
    return hipGetLastError();
}
template hipError_t launch_interp_atmos<double>(const GridDev&, const AtmosDev<double>&, const double*,
                                                const double*, const ne_exchange_state&, hipStream_t);
template hipError_t launch_interp_atmos<float>(const GridDev&, const AtmosDev<float>&, const double*,
                                               const double*, const ne_exchange_state&, hipStream_t);

template <typename Real>
hipError_t launch_interp_radiation(const GridDev& g, const RadDev<Real>& r, const double* fi, const double* fj,
                                   View3<double> sw, View3<double> lw, hipStream_t s) {
    hipLaunchKernelGGL(interp_radiation_kernel<Real>, grid2d(g.Nx, g.Ny), dim3(BX, BY), 0, s, g, r, fi, fj, sw, lw);
    if (g.ring)
        hipLaunchKernelGGL(interp_radiation_ring_kernel<Real>, dim3((ring_size(g) + 255) / 256), dim3(256), 0, s, g,
                           r, fi, fj, sw, lw);
    return hipGetLastError();
}
template hipError_t launch_interp_radiation<double>(const GridDev&, const RadDev<double>&, const double*,
                                                    const double*, View3<double>, View3<double>, hipStream_t);
template hipError_t launch_interp_radiation<float>(const GridDev&, const RadDev<float>&, const double*,
                                                   const double*, View3<double>, View3<double>, hipStream_t);

hipError_t launch_flux_solve_exchange(const GridDev& g, const ne_params& p, const ne_exchange_state& a,
                                      const ne_ocean_state& o, const FluxOut& out, ne_counters* c,
                                      hipStream_t s) {
    View3<const double> sw = view<const double>(a.downwelling_shortwave);
    View3<const double> lw = view<const double>(a.downwelling_longwave);
    const int path = flux_path(p);
    auto kernel = path == 2 ? flux_solve_exchange_kernel<true, true>
                : path == 1 ? flux_solve_exchange_kernel<true, false> : flux_solve_exchange_kernel<false, false>;
    hipLaunchKernelGGL(kernel, grid2d(g.Nx, g.Ny), dim3(BX, BY), 0, s, g, p, view<const double>(a.u),
                       view<const double>(a.v), view<const double>(a.T), view<const double>(a.p),
                       view<const double>(a.q), view<const double>(o.T), view<const double>(o.S),
                       view<const double>(o.u), view<const double>(o.v), sw, lw, out, c);
    if (g.ring) {
        auto ring = path == 2 ? flux_solve_ring_kernel<true, true>
                  : path == 1 ? flux_solve_ring_kernel<true, false> : flux_solve_ring_kernel<false, false>;
        hipLaunchKernelGGL(ring, dim3((ring_size(g) + BX * BY - 1) / (BX * BY)), dim3(BX * BY), 0, s, g, p,
                           view<const double>(a.u), view<const double>(a.v), view<const double>(a.T),
                           view<const double>(a.p), view<const double>(a.q), view<const double>(o.T),
                           view<const double>(o.S), view<const double>(o.u), view<const double>(o.v), sw, lw, out);
    }
    return hipGetLastError();
}

hipError_t launch_freeze_clamp(const GridDev& g, const ne_params& p, View3<double> T, View3<const double> S,
                               View3<double> frazil, double dt, hipStream_t s) {
    hipLaunchKernelGGL(freeze_clamp_kernel, grid2d(g.Nx, g.Ny), dim3(BX, BY), 0, s, g, p, T, S, frazil, dt);
    return hipGetLastError();
}

hipError_t launch_assemble(const GridDev& g, const ne_params& p, const ne_interface_fluxes& f,
                           const ne_exchange_state& a, View3<const double> sst, const NetOut& n, hipStream_t s) {
    hipLaunchKernelGGL(assemble_kernel, grid2d(g.Nx, g.Ny), dim3(BX, BY), 0, s, g, p, view<double>(f.latent_heat),
                       view<double>(f.sensible_heat), view<double>(f.water_vapor), view<double>(f.x_momentum),
                       view<double>(f.y_momentum), view<double>(a.rain), view<double>(a.snow), sst, n);
    return hipGetLastError();
}

hipError_t launch_radiation(const GridDev& g, const ne_params& p, View3<const double> sw, View3<const double> lw,
                            View3<const double> Ts, const NetOut& n, const RadOut& r, hipStream_t s) {
    hipLaunchKernelGGL(radiation_kernel, grid2d(g.Nx, g.Ny), dim3(BX, BY), 0, s, g, p, sw, lw, Ts, n, r);
    return hipGetLastError();
}

// The parameters for which the FP32 form's flux solve runs its FP32 iteration (flux_column,
// Mixed && Fast): every column then either finishes in FP32 or is deferred.
bool mixed_iteration_applies(const ne_params& p) {
    const bool shared_scalar = p.temperature_roughness == p.water_vapor_roughness &&
                               (p.temperature_roughness == NE_ROUGHNESS_EDSON ||
                                p.temperature_roughness_length == p.water_vapor_roughness_length);
    return NE_MIXED_PREITERATION && p.stop_criteria == NE_STOP_CONVERGENCE && shared_scalar &&
           p.momentum_roughness == NE_ROUGHNESS_EDSON && p.water_vapor_roughness == NE_ROUGHNESS_EDSON;
}

template <typename Real>
hipError_t launch_surface_update(const GridDev& g, const ne_params& p, const AtmosDev<Real>& a, const double* fi,
                                 const double* fj, const RadDev<Real>& r, const double* rfi, const double* rfj,
                                 View3<Real> T, View3<const Real> S, View3<const Real> u, View3<const Real> v,
                                 const FluxOut& out, const NetOut& n, const RadOut& ro, View3<double> frazil,
                                 double dt_frazil, const uint8_t* freeze_flags, ne_counters* c, int* deferred,
                                 hipStream_t s) {
    const int path = flux_path(p);
    const dim3 grid = grid2d(g.Nx, g.Ny);
    if (sizeof(Real) == 4 && path >= 1 && deferred && mixed_iteration_applies(p)) {
        // the FP32 iteration alone at 6 wavefronts per SIMD in blocks of 8 rows, then the deferred
        // columns
        constexpr int KBY = NE_FLUX_MIXED_ROWS;
        auto k1 = path == 2 ? surface_update_kernel<Real, true, true, true, KBY>
                            : surface_update_kernel<Real, true, false, true, KBY>;
        auto k2 = path == 2 ? surface_deferred_kernel<Real, true, true> : surface_deferred_kernel<Real, true, false>;
        hipLaunchKernelGGL(k1, dim3((g.Nx + BX - 1) / BX, (g.Ny + KBY - 1) / KBY), dim3(BX, KBY), 0, s, g, p, a, fi, fj,
                           r, rfi, rfj, T, S, u, v, out, n, ro, frazil, dt_frazil, freeze_flags, c, deferred);
        const unsigned blocks = std::min<unsigned>(grid.x * grid.y, 256u);
        hipLaunchKernelGGL(k2, dim3(blocks), dim3(BX, BY), 0, s, g, p, a, fi, fj, r, rfi, rfj, T, S, u, v, out, n, ro,
                           frazil, dt_frazil, freeze_flags, c, deferred);
        return hipGetLastError();
    }
    auto kernel = path == 2 ? surface_update_kernel<Real, true, true>
                : path == 1 ? surface_update_kernel<Real, true, false> : surface_update_kernel<Real, false, false>;
    hipLaunchKernelGGL(kernel, grid, dim3(BX, BY), 0, s, g, p, a, fi, fj, r, rfi, rfj, T, S, u, v, out, n, ro, frazil,
                       dt_frazil, freeze_flags, c, (int*)nullptr);
    return hipGetLastError();
}
template hipError_t launch_surface_update<double>(const GridDev&, const ne_params&, const AtmosDev<double>&,
                                                  const double*, const double*, const RadDev<double>&, const double*,
                                                  const double*, View3<double>, View3<const double>,
                                                  View3<const double>, View3<const double>, const FluxOut&,
                                                  const NetOut&, const RadOut&, View3<double>, double,
                                                  const uint8_t*, ne_counters*, int*, hipStream_t);
template hipError_t launch_surface_update<float>(const GridDev&, const ne_params&, const AtmosDev<float>&,
                                                 const double*, const double*, const RadDev<float>&, const double*,
                                                 const double*, View3<float>, View3<const float>,
                                                 View3<const float>, View3<const float>, const FluxOut&,
                                                 const NetOut&, const RadOut&, View3<double>, double,
                                                 const uint8_t*, ne_counters*, int*, hipStream_t);

hipError_t launch_surface_tau(const GridDev& g, const ne_params& p, const FluxOut& f, const NetOut& n,
                              int* deferred, hipStream_t s) {
    hipLaunchKernelGGL(surface_tau_kernel, grid2d(g.Nx, g.Ny), dim3(BX, BY), 0, s, g, p, f, n, deferred);
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_batch_surface_solve(const BatchRec<Real>* recs, const WorkItem* work, int nwork, int block,
                                      int blocks, int path, bool fused, double dt_frazil, ne_counters* c,
                                      hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    if (block <= 0 || block > BX * BY || (block & (block - 1)) != 0 || blocks <= 0) return hipErrorInvalidValue;
    auto pick = [&](auto fused_tag) {
        constexpr bool F = decltype(fused_tag)::value;
        return path == 2 ? batch_surface_solve_kernel<Real, true, true, F>
             : path == 1 ? batch_surface_solve_kernel<Real, true, false, F> : batch_surface_solve_kernel<Real, false, false, F>;
    };
    auto kernel = fused ? pick(std::true_type{}) : pick(std::false_type{});
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(block), 0, s, recs, work, nwork, dt_frazil, c);
    return hipGetLastError();
}
template hipError_t launch_batch_surface_solve<double>(const BatchRec<double>*, const WorkItem*, int, int, int, int,
                                                       bool, double, ne_counters*, hipStream_t);
template hipError_t launch_batch_surface_solve<float>(const BatchRec<float>*, const WorkItem*, int, int, int, int,
                                                      bool, double, ne_counters*, hipStream_t);

template <typename Real>
hipError_t launch_batch_surface_tau(const BatchRec<Real>* recs, const WorkItem* work, int nwork, int block,
                                   int blocks, hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    if (block <= 0 || block > BX * BY || blocks <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_surface_tau_kernel<Real>, dim3(blocks), dim3(block), 0, s, recs, work, nwork);
    return hipGetLastError();
}
template hipError_t launch_batch_surface_tau<double>(const BatchRec<double>*, const WorkItem*, int, int, int,
                                                     hipStream_t);
template hipError_t launch_batch_surface_tau<float>(const BatchRec<float>*, const WorkItem*, int, int, int,
                                                    hipStream_t);

}  // namespace ne
