// ne_tracer.hip — WENO5 tracer-advection tendency on the immersed LatLon grid, fused
// with the top flux boundary condition and the split Runge-Kutta stage update.
//
// One work-item per column (i, j), marching up the active levels; blocks of 64×4
// with i fastest, so every level of every field is read by 64 consecutive lanes.
// The vertical face flux computed at the top face of level k is carried in
// registers as the bottom face of level k+1 (each z-face reconstructed once).
#include "ne_internal.h"

namespace ne {

namespace {

constexpr int BX = 64;
constexpr int BY = 4;
constexpr int KMIN_OUTSIDE = 1 << 30;  // kmin of a column outside a bounded domain

inline dim3 grid2d(int Nx, int Ny) { return dim3((Nx + BX - 1) / BX, (Ny + BY - 1) / BY); }

// Advective flux through one face: A·vel·c̃ (c̃ from the six cells around the face).
template <typename R>
__device__ __forceinline__ R face_flux(unsigned act6, const R c[6], R vel, R A) {
    int order = face_order(act6);
    if (order == 0) return (R)0;
    return A * vel * face_value(c, order, vel);
}

// MODE 0: c_out = c0 + beta_dt·G(c_in) on active cells. MODE 1: write G (0 on inactive cells).
template <typename Real, int MODE>
__global__ void __launch_bounds__(BX* BY)
tracer_kernel(GridDev g, View3<const Real> u, View3<const Real> v, View3<const Real> w, View3<const Real> T0,
              View3<const Real> S0, View3<const Real> Tin, View3<const Real> Sin, View3<Real> Tout,
              View3<Real> Sout, View3<const double> JT, View3<const double> JS, View3<const double> Jw,
              double beta_dt) {
    typedef Real R;
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    if (i >= g.Nx || j >= g.Ny) return;
    const int Nz = g.Nz;
    const int kmin_c = g.kmin[i + (int64_t)j * g.Nx];
    if (MODE == 1) {
        for (int k = 0; k < (kmin_c < Nz ? kmin_c : Nz); k++) { Tout(i, j, k) = (R)0; Sout(i, j, k) = (R)0; }
    }
    if (kmin_c >= Nz) return;

    // kmin of the neighbour columns i-3 … i+3 and j-3 … j+3 (outside a bounded domain: never active).
    int kx[7], ky[7], ix[7], jy[7];
#pragma unroll
    for (int n = 0; n < 7; n++) {
        int64_t cx = g.column(i - 3 + n, j);
        kx[n] = cx >= 0 ? g.kmin[cx] : KMIN_OUTSIDE;
        ix[n] = g.wx(i - 3 + n);
        int64_t cy = g.column(i, j - 3 + n);
        ky[n] = cy >= 0 ? g.kmin[cy] : KMIN_OUTSIDE;
        jy[n] = g.wy(j - 3 + n);
    }
    const int iE = g.wx(i + 1);
    const int jN = g.wy(j + 1);
    const double Ayf_s = g.dx_yface[jy[3]], Ayf_n = g.dx_yface[jN];
    const double Axf = g.dy_xface[j];
    const double Azc = g.Az[j];

    R FzT = (R)0, FzS = (R)0;  // flux through the bottom face of the current level
    for (int k = kmin_c; k < Nz; k++) {
        R div_T = (R)0, div_S = (R)0;
        if (g.advection) {
            const R dzk = (R)g.dz[k];
            // x faces: cells i-3 … i+3
            unsigned ax = 0;
            R cT[7], cS[7];
#pragma unroll
            for (int n = 0; n < 7; n++) {
                bool a = k >= kx[n];
                ax |= (a ? 1u : 0u) << n;
                cT[n] = a ? Tin(ix[n], j, k) : (R)0;
                cS[n] = a ? Sin(ix[n], j, k) : (R)0;
            }
            const R Ax = (R)Axf * dzk;
            R uw = (ax & 0x0Cu) == 0x0Cu ? u(ix[3], j, k) : (R)0;
            R ue = (ax & 0x18u) == 0x18u ? u(iE, j, k) : (R)0;
            R FxwT = face_flux(ax & 0x3Fu, cT, uw, Ax), FxeT = face_flux((ax >> 1) & 0x3Fu, cT + 1, ue, Ax);
            R FxwS = face_flux(ax & 0x3Fu, cS, uw, Ax), FxeS = face_flux((ax >> 1) & 0x3Fu, cS + 1, ue, Ax);
            // y faces: cells j-3 … j+3
            unsigned ay = 0;
#pragma unroll
            for (int n = 0; n < 7; n++) {
                bool a = k >= ky[n];
                ay |= (a ? 1u : 0u) << n;
                cT[n] = a ? Tin(i, jy[n], k) : (R)0;
                cS[n] = a ? Sin(i, jy[n], k) : (R)0;
            }
            const R Ays = (R)Ayf_s * dzk, Ayn = (R)Ayf_n * dzk;
            R vs = (ay & 0x0Cu) == 0x0Cu ? v(i, jy[3], k) : (R)0;
            R vn = (ay & 0x18u) == 0x18u ? v(i, jN, k) : (R)0;
            R FysT = face_flux(ay & 0x3Fu, cT, vs, Ays), FynT = face_flux((ay >> 1) & 0x3Fu, cT + 1, vn, Ayn);
            R FysS = face_flux(ay & 0x3Fu, cS, vs, Ays), FynS = face_flux((ay >> 1) & 0x3Fu, cS + 1, vn, Ayn);
            // z face k+1 (top of this level): cells k-2 … k+3 of this column
            unsigned az = 0;
            R zT[6], zS[6];
#pragma unroll
            for (int n = 0; n < 6; n++) {
                int kk = k - 2 + n;
                bool a = kk >= kmin_c && kk < Nz;
                az |= (a ? 1u : 0u) << n;
                zT[n] = a ? Tin(i, j, kk) : (R)0;
                zS[n] = a ? Sin(i, j, kk) : (R)0;
            }
            R wt = (az & 0x0Cu) == 0x0Cu ? w(i, j, k + 1) : (R)0;
            R FztT = face_flux(az, zT, wt, (R)Azc), FztS = face_flux(az, zS, wt, (R)Azc);
            div_T = (FxeT - FxwT) + (FynT - FysT);
            div_T = div_T + (FztT - FzT);
            div_S = (FxeS - FxwS) + (FynS - FysS);
            div_S = div_S + (FztS - FzS);
            FzT = FztT;
            FzS = FztS;
        }
        const R V = (R)(Azc * g.dz[k]);
        R GT = -((R)1 / V * div_T);
        R GS = -((R)1 / V * div_S);
        if (k == Nz - 1) {
            // MultipleFluxes(Jᵀ, FreshwaterExchange{:T}) = Jᵀ + 0 on a static grid;
            // MultipleFluxes(Jˢ, FreshwaterExchange{:S}) = Jˢ + (Sᴺ Jʷ − 0)
            // (multiple_surface_fluxes.jl:27-31; ocean_simulation.jl:80-91)
            double FT = JT.at2(i, j) + (0.0 + 0.0);
            double FS = JS.at2(i, j) + (((double)Sin(i, j, k) * Jw.at2(i, j) - 0.0) + 0.0);
            GT = GT - (R)(FT * Azc) / V;
            GS = GS - (R)(FS * Azc) / V;
        }
        if (MODE == 0) {
            Tout(i, j, k) = T0(i, j, k) + (R)beta_dt * GT;
            Sout(i, j, k) = S0(i, j, k) + (R)beta_dt * GS;
        } else {
            Tout(i, j, k) = GT;
            Sout(i, j, k) = GS;
        }
    }
}

// ---------------------------------------------------------------------------
// Compensated budget: Neumaier per column, then a fixed-order double-double tree.
// ---------------------------------------------------------------------------

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
    hipLaunchKernelGGL((tracer_kernel<Real, 0>), grid2d(g.Nx, g.Ny), dim3(BX, BY), 0, s, g, u, v, w, T0, S0, Tin,
                       Sin, Tout, Sout, JT, JS, Jw, beta_dt);
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
    hipLaunchKernelGGL((tracer_kernel<double, 1>), grid2d(g.Nx, g.Ny), dim3(BX, BY), 0, s, g, u, v, w, T, S, T, S,
                       GT, GS, JT, JS, Jw, 0.0);
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

}  // namespace ne
