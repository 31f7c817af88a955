// Declarations shared by the boundary/source kernels and the C ABI.
#pragma once
#include <hip/hip_runtime.h>

namespace magic {

template <typename T>
__device__ __forceinline__ T fminD(T a, T b) { return a < b ? a : b; }
template <typename T>
__device__ __forceinline__ T fmaxD(T a, T b) { return a > b ? a : b; }

struct Dims {
    int mx, my, mz, mbc;
    int NX, NY, NZ;       // padded extents
    long long P;          // NX*NY*NZ
};

// Background planes the source pass reads (aux slots 1-5, 18, 19).
struct Aux {
    const double *rho, *mx, *my, *mz, *E, *R, *gam;
};

// Forcing of src3 (forcemth 1 and 3), evaluated at time t over dt.
struct Forcing {
    int forcemth;
    double t, dt, xlower, ylower, zlower, dx, dy, dz;
    double amplitude, xpos, ypos, zpos, xwidth, ywidth, zwidth, tcenter, twidth;
    double omega, propx, propz, vsrcx;
    double tgauss;        // exp(-.5*((t-tcenter)**2/twidth**2)), host FP64
};

// Gate of the shared kernels for a single grid: always on; kernel
// arguments are used as passed.
struct NoGate {
    __device__ __forceinline__ bool on() const { return true; }
    template <typename T>
    __device__ __forceinline__ T f(T x, int) const { return x; }
    __device__ __forceinline__ void forcing(double &, double &, double &) const {}
};

template <typename T, class G>
__global__ void k_bc_x(Dims d, T *q, int nc, int lo, int hi, G gate);
template <typename T, class G>
__global__ void k_bc_y(Dims d, T *q, int nc, int lo, int hi, G gate);
// Stephan & Alexander convective heating profile (src3.f90:534-572) for a
// precipitation rate precipw > 0.1 at height zcell (m): K/s.
__device__ __forceinline__ double sa_heating(double precipw, double zcell)
{
    const double pi = 3.141592653589793;
    const double th = 10.79 + 0.05 * precipw, ph = 6.93 - 0.03 * precipw;
    const double hamp = -0.0001 + 0.0025 * precipw;
    const double tc = 1.12 + 0.10 * precipw, pc = 0.63 + 0.02 * precipw;
    const double bc = 0.37 + 0.00 * precipw, camp = -0.0003 - 0.0005 * precipw;
    const double z2 = zcell / 1000.0;
    double heat;
    if ((bc <= z2) && (z2 <= pc)) heat = camp * sin((2.0 * pi / (4.0 * (pc - bc))) * (z2 - bc));
    else if ((pc <= z2) && (z2 <= tc)) heat = -camp * sin((2.0 * pi / (4.0 * (tc - pc))) * (z2 - tc));
    else if ((pc <= z2) && (z2 <= ph)) heat = hamp * sin((2.0 * pi / (4.0 * (ph - tc))) * (z2 - tc));
    else if ((ph <= z2) && (z2 <= th)) heat = -hamp * sin((2.0 * pi / (4.0 * (th - ph))) * (z2 - th));
    else heat = 0.0;
    return heat;
}

template <class G>
__global__ void k_bc_z(Dims d, double *q, int nc, int lo, int hi, const double *aux1,
                       const double *aux2, const double *aux3, const double *aux5,
                       const double *zratio, G gate);
template <typename T, class G>
__global__ void k_clamp(T *q, long long P, int nc, G gate);
template <bool COL, class G>
__global__ void k_src_convert(Dims d, double *q, const Aux a, G gate, const double *dq, int nc);
template <typename T, class G>
__global__ void k_zghost_uvw(Dims d, T *q, G gate);
template <typename T, bool COLC, class G>
__global__ void k_diff_cross(Dims d, T *u, const T *q, const T *coef, T dxy, T dxz, T dyz, int kb, G gate);
template <typename T, bool COLC, class G>
__global__ void k_diff_xyz(Dims d, T *q, const T *u, const T *coef, T fx, T fy, T fz, int kb, G gate);
template <typename T, class G>
__global__ void k_copy_interior(Dims d, T *dst, const T *src, int nplanes, G gate);
template <typename T, class G>
__global__ void k_conduct_zghost(Dims d, T *u, G gate);
template <typename T, class G>
__global__ void k_conduct_zspill(Dims d, T *u, int S, G gate);
template <typename T, bool COLC, class G>
__global__ void k_pass(Dims d, T *out, const T *src, const T *ghost, const T *coef, T f, int axis, G gate);
template <bool COL, class G>
__global__ void k_src_finish(Dims d, double *q, int nc, const Aux a, Forcing f, G gate);
__global__ void k_aos_to_planes(double *planes, const double *aos, long long P, int nc);

// ---- FP32 perturbation form (bc_src32.hip) ----
// Background columns at rest, indexed by k + mbc - 1: rho_bar, E_bar,
// gamma (aux19), R (aux18), and rho_bar, E_bar in FP64.
struct Col32 {
    const double *rho, *E, *gam, *R;   // FP64 columns of rho_bar, E_bar, gamma, R
};
// bc3 z-face factors (rho_bar_ghost/rho_bar_boundary)**0.5 per side and layer
struct Zr32 {
    double s[2][8];
};
template <class G>
__global__ void k32_bc_z(Dims d, float *q, int nc, int lo, int hi, Zr32 zr, G gate);
template <class G>
__global__ void k32_src_convert(Dims d, float *q, Col32 col, G gate, const float *dq, int nc);
template <class G>
__global__ void k32_src_finish(Dims d, float *q, int nc, Col32 col, Forcing f, G gate);
template <typename T, class G>
__global__ void k_copy(T *dst, const T *src, long long n, G gate);
__global__ void k32_aos_to_pert(Dims d, float *q, const double *aos, int nc, const double *bgrho,
                                const double *bgE);
__global__ void k32_pert_to_aos(Dims d, double *aos, const float *q, int nc, const double *bgrho,
                                const double *bgE);
__global__ void k_planes_to_aos(double *aos, const double *planes, long long P, int nc);

}  // namespace magic
