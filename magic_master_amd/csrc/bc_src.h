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

template <typename T>
__global__ void k_bc_x(Dims d, T *q, int nc, int lo, int hi);
template <typename T>
__global__ void k_bc_y(Dims d, T *q, int nc, int lo, int hi);
__global__ void k_bc_z(Dims d, double *q, int nc, int lo, int hi, const double *aux1,
                       const double *aux2, const double *aux3, const double *aux5,
                       const double *zratio);
template <typename T>
__global__ void k_clamp(T *q, long long P, int nc);
__global__ void k_src_convert(long long P, double *q, const Aux a);
template <typename T>
__global__ void k_zghost_uvw(Dims d, T *q);
template <typename T, bool COLC>
__global__ void k_diff_cross(Dims d, T *q, const T *u, const T *coef, T dxy, T dxz, T dyz);
template <typename T>
__global__ void k_copy_interior(Dims d, T *dst, const T *src, int nplanes);
template <typename T>
__global__ void k_conduct_zghost(Dims d, T *u);
template <typename T>
__global__ void k_conduct_zspill(Dims d, T *u, int S);
template <typename T, bool COLC>
__global__ void k_pass(Dims d, T *out, const T *src, const T *ghost, const T *coef, T f, int axis);
__global__ void k_src_finish(Dims d, double *q, int nc, const Aux a, Forcing f);
__global__ void k_aos_to_planes(double *planes, const double *aos, long long P, int nc);
__global__ void k_planes_to_aos(double *aos, const double *planes, long long P, int nc);

}  // namespace magic
