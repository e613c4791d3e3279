// loopctl.hip -- the Green loop's per-iteration control on the device.
//
// After each iteration's Newton launch and the first stage of its partial-sum
// reduction, k_loop_ctl (one workgroup) finishes the reduction, then does
// what the host did between iterations: the Newton status and finiteness
// checks, get_smacro (src/evolution.c:473-547, with the mixed boundary
// condition's macro strain correction dD), the normalised error norms
// (src/evolution.c:1449-1455) and the loop test (src/evolution.c:1354).  The
// next iteration's strain update reads dD from the control block, so the
// host never has to wait for an iteration to finish before queueing the
// next; it reads the control flags once per batch of iterations.  Once a
// phase of the loop has ended the flag `done` is set and every kernel of the
// iterations queued after it returns at once.
//
// The arithmetic is the host code's it replaces, operation for operation
// (built with -ffp-contract=off; IEEE division and square root), so the
// stopping decisions and every value are the same; get_smacro's part lives in
// loopctl_step.h, whose host build equals the oracle's bit for bit.
#include <math.h>

#include "evpgpu_impl.h"

namespace {

__global__ void k_ctl_init(LoopCtl *L, LoopCtl init, CtlConst *Kd, CtlConst K)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    *L = init;
    *Kd = K;
}

// a new phase of the loop: first != 0 runs at least one iteration (the
// do-while of the FP64 sweeps); an error that ended the last phase stays
__global__ void k_ctl_phase(LoopCtl *Lp, int limit, int conv, int first)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    LoopCtl L = *Lp;
    L.limit = limit;
    L.conv = conv;
    L.f.done = L.f.status ? 1 : (first ? 0 : lctl::phase_done(L));
    Lp->limit = limit;
    Lp->conv = conv;
    Lp->f.done = L.f.done;
}

constexpr int NL = (int)(sizeof(LoopCtl) / sizeof(double));
constexpr int NK = (int)(sizeof(CtlConst) / sizeof(double));
static_assert(sizeof(LoopCtl) % sizeof(double) == 0 && sizeof(CtlConst) % sizeof(double) == 0,
              "control blocks are whole doubles");
static_assert(NL + NK <= 256, "one double per thread");

// Stage 2 of the Newton partial sums (k_reduce_rows' order: thread t sums
// value t % 16 over the stage-1 rows t / 16, t / 16 + 16, ..., then the 16
// row lanes of each value are summed in order), then the control step.  The
// control block and the constants come in through LDS (one double per
// thread, loaded together with the partial sums): one trip to memory.
__global__ __launch_bounds__(256) void k_loop_ctl(const double *__restrict__ red2, int g,
                                                  const int *__restrict__ status, LoopCtl *Lp,
                                                  const CtlConst *__restrict__ Kp)
{
    __shared__ double sh[256];
    __shared__ double red[NRED];
    __shared__ double blk[NL + NK];
    const int t = threadIdx.x;
    if (t < NL) blk[t] = reinterpret_cast<const double *>(Lp)[t];
    else if (t < NL + NK) blk[t] = reinterpret_cast<const double *>(Kp)[t - NL];
    int st = t == 0 ? *status : 0;
    const int v = t % NRED, r0 = t / NRED;
    // every row this thread sums is loaded before the first add (one trip to
    // memory: the rows were written by workgroups on all the XCDs), then
    // summed in row order
    constexpr int STEP = 256 / NRED, TRIPS = (REDUCE_G + STEP - 1) / STEP;
    double x[TRIPS];
#pragma unroll
    for (int k = 0; k < TRIPS; k++) x[k] = r0 + k * STEP < g ? red2[(r0 + k * STEP) * NRED + v] : 0.0;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < TRIPS; k++)
        if (r0 + k * STEP < g) s += x[k];
    sh[t] = s;
    __syncthreads();
    const LoopCtl &Ls = *reinterpret_cast<const LoopCtl *>(blk);
    if (Ls.f.done) return;
    if (t < NRED) {
        double u = 0.0;
#pragma unroll
        for (int r = 0; r < 256 / NRED; r++) u += sh[r * NRED + t];
        red[t] = u;
    }
    __syncthreads();
    if (t != 0) return;
    LoopCtl L = Ls;
    const CtlConst &K = *reinterpret_cast<const CtlConst *>(blk + NL);
    // the strain update of this iteration added the dD of the last one
#pragma unroll
    for (int i = 0; i < 9; i++) L.dD_applied[i] += L.dD[i];
    if (!st) {
        // a non-finite rate makes the block sums non-finite; the reference
        // stops on such values (assert(!isnan), src/constitutive.c:617,653,700)
#pragma unroll
        for (int k = 0; k < NRED; k++)
            if (!isfinite(red[k])) st = EVPGPU_ERR_NONFINITE;
    }
    if (st) {
        Lp->f.status = st;
        Lp->f.done = 1;
        return;
    }
    L.err_e = red[1] * K.scale;
    L.err_s = red[0] * K.scale;
    L.nit += red[2];
    L.ntr += red[15];
    lctl::smacro(K, L, red);
    if (fabs(L.e_vm) < 1E-5) L.err_e = 0.0;
    else L.err_e /= L.e_vm;
    L.err_s /= L.s_vm;
    const double *d = L.dD;
    L.add[0] = d[0];
    L.add[1] = d[4];
    L.add[2] = d[8];
    L.add[3] = (d[5] + d[7]) / 2.0;
    L.add[4] = (d[2] + d[6]) / 2.0;
    L.add[5] = (d[1] + d[3]) / 2.0;
    L.f.iter++;
    L.f.done = lctl::phase_done(L);
    *Lp = L;
}

} // namespace

int ctl_init_launch(evpgpu_ctx *c, const LoopCtl &init, const CtlConst &K)
{
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, c->stream, c->ctl, init, c->ctl_k, K);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int ctl_phase_launch(evpgpu_ctx *c, int limit, int conv, int first)
{
    hipLaunchKernelGGL(k_ctl_phase, dim3(1), dim3(64), 0, c->stream, c->ctl, limit, conv, first);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}

int ctl_iter_launch(evpgpu_ctx *c, int nblk)
{
    int g = 0;
    int st = reduce_stage1(c, c->partials, nblk, NRED, c->skip, &g);
    if (st) return st;
    if (g < 1 || g > REDUCE_G) return EVPGPU_ERR_ARG; // k_loop_ctl loads at most REDUCE_G rows
    hipLaunchKernelGGL(k_loop_ctl, dim3(1), dim3(256), 0, c->stream, (const double *)c->red2, g,
                       (const int *)c->d_status, c->ctl, (const CtlConst *)c->ctl_k);
    return hipGetLastError() == hipSuccess ? EVPGPU_OK : EVPGPU_ERR_HIP;
}
