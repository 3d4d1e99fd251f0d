// pfc_noalloc.hip -- pfc_smooth_noalloc: the Gaussian smoothing of pfc_smooth
// on caller-owned device buffers, with every work buffer in one static
// device array, so that the call allocates nothing (for callers under HIP
// graph capture, serving loops and harnesses that forbid allocation).
//
// Work memory, carved from `scratch` at 256-byte boundaries:
//   twx   Nx complex twiddles  e^{-2 pi i m/Nx}
//   twy   Ny complex twiddles  e^{-2 pi i m/Ny}
//   spec  the packed row transforms (column-major, padded pitch), smoothed
//         in place along y
//   stage (mixed form only) Nx*Ny FP32 reals: the input and then the output
//         in FP32
// The array is a host-side stand-in in host code: its device address is
// looked up once per device with hipGetSymbolAddress and only that address
// is handed to kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/pfc_relax.h"
#include "pfc_launch.hpp"

using namespace pfc;

static __device__ __align__(256) unsigned char scratch[PFC_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices];   // device address of `scratch` per device

// The current device's address of `scratch`, looked up on the first call on
// that device (with that device current).
int scratch_ptr(unsigned char** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return PFC_EDEVICE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PFC_EDEVICE;
        g_scratch[dev] = p;
    }
    *out = (unsigned char*)g_scratch[dev];
    return PFC_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// tw[m] = e^{-2 pi i m/n}, from the FP64 sincospi of -2m/n
template <typename T>
__global__ void k_twiddles(cpx<T>* tw, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) {
        double s, c;
        sincospi(-2.0 * (double)m / (double)n, &s, &c);
        tw[m] = {(T)c, (T)s};
    }
}

__global__ void k_to_f32(const double* __restrict__ in, float* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}
__global__ void k_to_f64(const float* __restrict__ in, double* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}

// Bytes of `scratch` one call needs (the layout at the top of this file).
size_t need_bytes(int Nx, int Ny, size_t esz, size_t nk, bool stage) {
    return align256((size_t)Nx * esz) + align256((size_t)Ny * esz) + align256(nk * esz) +
           (stage ? align256((size_t)Nx * Ny * sizeof(float)) : 0);
}

template <typename T>
int smooth(const double* in, double* out, int Nx, int Ny, double dx, double dy, double sigma,
           hipStream_t st) {
    constexpr bool F32 = sizeof(T) == 4;
    constexpr size_t esz = sizeof(cpx<T>);
    pfc_ctx c;   // launch parameters only: no buffer of its own
    std::memset(&c, 0, sizeof c);
    c.Nx = Nx;
    c.Ny = Ny;
    c.U = 1;
    c.prec = F32 ? PFC_MIXED : PFC_FP64;
    c.esz = esz;
    c.nr = (size_t)Nx * Ny;
    c.sC = c.sI = c.sJ = ilog2(Ny);
    c.nk = lay_elems(c.sC, Nx / 2, Ny, (unsigned)esz);
    c.stream = st;
    const double pi = 3.14159265358979323846264338327;
    c.hp.dkx = 2.0 * pi / (dx * Nx);
    c.hp.dky = 2.0 * pi / (dy * Ny);
    if (need_bytes(Nx, Ny, esz, c.nk, F32) > (size_t)PFC_NOALLOC_SCRATCH_BYTES) return PFC_ESIZE;
    unsigned char* base = nullptr;
    const int rc = scratch_ptr(&base);
    if (rc) return rc;
    size_t off = 0;
    c.twx = base + off;
    off += align256((size_t)Nx * esz);
    c.twy = base + off;
    off += align256((size_t)Ny * esz);
    void* spec = base + off;
    off += align256(c.nk * esz);
    float* stage = F32 ? (float*)(base + off) : nullptr;

    hipLaunchKernelGGL(k_twiddles<T>, dim3((Nx + 255) / 256), dim3(256), 0, st, (cpx<T>*)c.twx, Nx);
    hipLaunchKernelGGL(k_twiddles<T>, dim3((Ny + 255) / 256), dim3(256), 0, st, (cpx<T>*)c.twy, Ny);
    const void* src = in;
    void* dst = out;
    if constexpr (F32) {
        hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st, in, stage, c.nr);
        src = stage;
        dst = stage;
    }
    Launch<T>::row_r2c_any(&c, src, spec, c.sC, 1);
    Launch<T>::col_smooth_any(&c, spec, spec, sigma);   // in place, column by column
    Launch<T>::row_c2r_any(&c, spec, c.sC, dst, nullptr, 1);
    if constexpr (F32) hipLaunchKernelGGL(k_to_f64, dim3(1024), dim3(256), 0, st, stage, out, c.nr);
    return hipGetLastError() == hipSuccess ? PFC_OK : PFC_EDEVICE;
}

}  // namespace

extern "C" int pfc_smooth_noalloc(const double* in, double* out, int Nx, int Ny, double dx, double dy,
                                  double sigma, int precision, void* stream) {
    if (!in || !out || !(sigma > 0.0) || !(dx > 0.0) || !(dy > 0.0)) return PFC_EINVAL;
    if (precision != PFC_FP64 && precision != PFC_MIXED) return PFC_EINVAL;
    if (!size_ok(Nx) || !size_ok(Ny)) return PFC_ESIZE;
    if ((long long)Nx * Ny > PFC_NOALLOC_MAX_POINTS) return PFC_ESIZE;
    hipStream_t st = (hipStream_t)stream;
    return precision == PFC_FP64 ? smooth<double>(in, out, Nx, Ny, dx, dy, sigma, st)
                                 : smooth<float>(in, out, Nx, Ny, dx, dy, sigma, st);
}
