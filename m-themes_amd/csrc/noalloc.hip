// noalloc.hip -- the per-voxel stress update on caller-owned device buffers
// with no device allocation: evpgpu_newton_device_noalloc.
//
// It runs the kernel evpgpu_newton_host runs (k_newton, the full update of
// src/evolution.c:44-310 for a batch of voxels) on device pointers the
// caller passes, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops).  The kernel's scratch -- NRED block partial
// sums per 64 voxels and the status word -- lives in the file-scope array
// `scratch` below, one copy per GPU, held until the process ends.  Its
// device address is taken once per device (hipGetSymbolAddress) on the first
// call made with that device current; host code passes only that address.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// race.
#include "evpgpu_impl.h"

namespace {

// largest batch: 2^21 voxels (a 128^3 grid), 32768 blocks of 16 partials
constexpr long long NOALLOC_MAX_N = EVPGPU_NOALLOC_MAX_VOXELS;
constexpr long long NOALLOC_NBLK = NOALLOC_MAX_N / 64;
// partial sums, then the status word (and its validation flag) in the last
// 256-byte line
constexpr long long SCRATCH_DOUBLES = NOALLOC_NBLK * NRED + 32;

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES];

// status = EVPGPU_ERR_ARG when a voxel's phase is outside 1..nphases or its
// orientation index outside 0..n_orient-1 (the kernel reads both unchecked)
__global__ void k_check_ids(const int *__restrict__ phase, const int *__restrict__ orient, long long n,
                            int nphases, int n_orient, int *status)
{
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int p = phase[v], o = orient[v];
    if (p < 1 || p > nphases || o < 0 || o >= n_orient) atomicCAS(status, 0, (int)EVPGPU_ERR_ARG);
}

int scratch_ptr(double **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return EVPGPU_ERR_HIP;
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return EVPGPU_ERR_HIP;
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return EVPGPU_OK;
}

} // namespace

extern "C" int evpgpu_newton_device_noalloc(evpgpu_ctx *c, int n, const int *phase, const int *orient,
                                            const double *sgr, int n_orient, double *sig6,
                                            const double *strain6, const double *eps6,
                                            const double *tranmat9, const double *crss24, double *edot6,
                                            double *gamdot12, double *err2, int *niter)
{
    if (!c || n < 1 || !phase || !orient || !sgr || n_orient < 1 || !sig6 || !strain6 || !eps6 ||
        !tranmat9 || !crss24 || !edot6 || !gamdot12 || !err2 || !niter || !c->have_phases ||
        !c->have_loading)
        return EVPGPU_ERR_ARG;
    if ((long long)n > NOALLOC_MAX_N) return EVPGPU_ERR_SIZE;
    double *sp = nullptr;
    int st = scratch_ptr(&sp);
    if (st) return st;
    double *partials = sp;
    int *status = (int *)(sp + NOALLOC_NBLK * NRED);
    // the status word starts at zero on every call (a fill kernel, not a
    // memset node)
    st = fill_zero_int(c, status, 1);
    if (st) return st;
    hipLaunchKernelGGL(k_check_ids, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, phase, orient,
                       (long long)n, c->nphases, n_orient, status);
    if (hipGetLastError() != hipSuccess) return EVPGPU_ERR_HIP;
    // the status comes back through the context's pinned word
    if (hipMemcpyAsync(c->h_status, status, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return EVPGPU_ERR_HIP;
    if (*c->h_status) return *c->h_status;
    NewtonParams P = context_newton_params(c);
    P.crss_sym = 0;
    P.dd = 0; // the power-law update on the caller's CRSS, as evpgpu_newton_host
    int nb = 0;
    st = newton_launch(c, P, sig6, sig6, strain6, eps6, edot6, tranmat9, crss24, gamdot12, phase, orient, sgr, n,
                       partials, &nb, status, err2, niter, 0);
    if (st) return st;
    if (hipMemcpyAsync(c->h_status, status, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return EVPGPU_ERR_HIP;
    return *c->h_status;
}
