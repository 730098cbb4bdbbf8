// hsh_adjoint.hip -- reverse-mode (adjoint) kernels of the hsolve step.
//
// One reverse step t (state V^t, slots S^t -> V^{t+1}) runs three kernels on
// the tile layout (block = one row of 64 neurons, lane = neuron):
//   k_adj_rhs    recomputes rhs/diag of step t from (V^t, S^t), as the
//                channel programs do (hines_chip.c:157-600), without writing
//                the slots;
//   k_adj_elim   the reverse of the Hines elimination (hines_solve.c:
//                246-295): it recomputes the eliminated diagonals/rhs and the
//                solution, then runs the back-substitution in reverse (rows
//                ascending, leaf to root) and the forward elimination in
//                reverse (rows descending) -- the transposed Hines sweep --
//                giving the adjoints of rhs, diag and of the three
//                coefficients of every row;
//   k_adj_local  the reverse of each compartment's channel program: from the
//                adjoints of rhs and diag, the adjoint of V^t, of every gate
//                state at t, and gradients of the program's constants (gmax,
//                Ek, Em/Rm, inject, tbyc, diagterm), accumulated over steps.
// Gate tables are piecewise linear in Vm: d(table)/dVm = (next row - row) *
// invdx inside the table, 0 where Vm is clamped.
#include "hsh_internal.h"
#include "hsh_device.h"
#include "hsh_adjoint_device.h"

namespace hsh {

namespace {

constexpr int ADJ_LDS_ROWS = 64;   // k_adj_elim_l: 4 x 64 x 64 doubles = 128 KiB of LDS

__global__ __launch_bounds__(256) void k_adj_rhs(AdjModel a)
{
    const DevModel& d = a.d;
    int b = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    b = __builtin_amdgcn_readfirstlane(b);
    if (b >= d.nblocks) return;
    const int lane = threadIdx.x & (W - 1);
    const BlockMeta bm = d.blocks[b];
    const long long pos = (long long)b * W + lane;
    const CompOut o = adj_forward_local(d, d.prog + bm.prog, a.St + bm.dslot, lane, a.Vt[pos]);
    a.RB[pos] = o.rhs;
    a.RD[pos] = o.diag;
}

__global__ __launch_bounds__(256) void k_adj_local(AdjModel a)
{
    const DevModel& d = a.d;
    int b = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    b = __builtin_amdgcn_readfirstlane(b);
    if (b >= d.nblocks) return;
    const int lane = threadIdx.x & (W - 1);
    const BlockMeta bm = d.blocks[b];
    const long long pos = (long long)b * W + lane;
    double vbar = adj_reverse_local(d, d.prog + bm.prog, a.St + bm.dslot, a.LS + bm.dslot, lane, a.Vt[pos],
                                    a.BB[pos], a.DB[pos]);
    // Crank-Nicolson: V^{t+1} = 2x - V^t
    if (d.crank) vbar -= a.LV[pos];
    a.LV[pos] = vbar;
}

// Reverse of the elimination, one lane per neuron, rows in global memory.
__global__ __launch_bounds__(256) void k_adj_elim(AdjModel a)
{
    const DevModel& d = a.d;
    int t = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= d.ntiles) return;
    const int lane = threadIdx.x & (W - 1);
    const TileMeta tm = d.tiles[t];
    const int R = tm.rows;
    const long long p0 = (long long)tm.block0 * W + lane;
#define Q(r) (p0 + (long long)(r) * W)
    double* Dp = a.Dp;
    double* Bp = a.Bp;
    double* X = a.X;
    double* XB = a.XB;
    double* BB = a.BB;
    double* DB = a.DB;
    double* Gta = a.GT;
    double* Gtf = a.GT + a.npos;
    double* Gtb = a.GT + 2 * a.npos;
    for (int r = 0; r < R; r++) {
        Dp[Q(r)] = a.RD[Q(r)];
        Bp[Q(r)] = a.RB[Q(r)];
    }
    // forward elimination and back-substitution, as the step ran them
    for (int r = 0; r < R - 1; r++) {
        const int p = d.par[Q(r)];
        if (p >= 0) {
            const double tt = d.ta[Q(r)] / Dp[Q(r)];
            Dp[Q(p)] -= d.tcf[Q(r)] * tt;
            Bp[Q(p)] -= Bp[Q(r)] * tt;
        }
    }
    for (int r = R - 1; r >= 0; r--) {
        const int p = d.par[Q(r)];
        double bv = Bp[Q(r)];
        if (p >= 0) bv -= d.tcb[Q(r)] * X[Q(p)];
        X[Q(r)] = bv / Dp[Q(r)];
    }
    const double xs = d.crank ? 2.0 : 1.0;
    for (int r = 0; r < R; r++) {
        XB[Q(r)] = xs * a.LV[Q(r)];
        BB[Q(r)] = 0.0;
        DB[Q(r)] = 0.0;
    }
    // back-substitution in reverse: rows ascending (children before parents)
    for (int r = 0; r < R; r++) {
        const int p = d.par[Q(r)];
        const double xb = XB[Q(r)];
        const double dk = Dp[Q(r)];
        BB[Q(r)] += xb / dk;
        DB[Q(r)] -= xb * X[Q(r)] / dk;
        if (p >= 0) {
            Gtb[Q(r)] -= xb * X[Q(p)] / dk;
            XB[Q(p)] -= xb * d.tcb[Q(r)] / dk;
        }
    }
    // forward elimination in reverse: rows descending
    for (int r = R - 2; r >= 0; r--) {
        const int p = d.par[Q(r)];
        if (p < 0) continue;
        const double dk = Dp[Q(r)];
        const double tt = d.ta[Q(r)] / dk;
        const double bbp = BB[Q(p)], dbp = DB[Q(p)];
        BB[Q(r)] -= bbp * tt;
        const double tbar = -bbp * Bp[Q(r)] - dbp * d.tcf[Q(r)];
        Gtf[Q(r)] -= dbp * tt;
        Gta[Q(r)] += tbar / dk;
        DB[Q(r)] -= tbar * tt / dk;
    }
#undef Q
}

// The same reverse sweep with the tile's working rows in LDS (one wave per
// tile, 4 x rows x 64 doubles): the row-to-row chain of the sweeps then waits
// on LDS instead of global memory, and what does not depend on the chain
// (tree coefficients, parent indices, the gradient accumulators' old values)
// is loaded ACH rows at a time into registers before those rows run.
// Arrays: eliminated diagonal, eliminated rhs, the solution (overwritten by
// the adjoint of diag once a row's reverse back-substitution is done: no
// later row reads it), the adjoint of the solution (overwritten by the
// adjoint of rhs, likewise).  Same operations in the same order as
// k_adj_elim; a zero-initialised accumulator of k_adj_elim is written here
// as 0.0 + v or 0.0 - v, so signed zeros agree as well.
constexpr int ACH = 8;

__global__ __launch_bounds__(64) void k_adj_elim_l(AdjModel a)
{
    extern __shared__ double lds[];
    const DevModel& d = a.d;
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    const TileMeta tm = d.tiles[t];
    const int R = tm.rows;
    const int M = a.max_rows;
    double* sD = lds;
    double* sB = lds + (size_t)M * W;
    double* sX = lds + (size_t)2 * M * W;
    double* sY = lds + (size_t)3 * M * W;
    const long long p0 = (long long)tm.block0 * W + lane;
#define Q(r) (p0 + (long long)(r) * W)
#define L(r) ((r) * W + lane)
    double* Gta = a.GT;
    double* Gtf = a.GT + a.npos;
    double* Gtb = a.GT + 2 * a.npos;
    const double xs = d.crank ? 2.0 : 1.0;
#pragma unroll 16
    for (int r = 0; r < R; r++) {
        sD[L(r)] = a.RD[Q(r)];
        sB[L(r)] = a.RB[Q(r)];
        sY[L(r)] = xs * a.LV[Q(r)];
    }
    // forward elimination and back-substitution, as the step ran them
    for (int r0 = 0; r0 < R - 1; r0 += ACH) {
        int pp[ACH];
        double ca[ACH], cf[ACH];
#pragma unroll
        for (int i = 0; i < ACH; i++) {
            const int r = r0 + i;
            pp[i] = -1;
            ca[i] = cf[i] = 0.0;
            if (r < R - 1) {
                pp[i] = d.par[Q(r)];
                ca[i] = d.ta[Q(r)];
                cf[i] = d.tcf[Q(r)];
            }
        }
#pragma unroll
        for (int i = 0; i < ACH; i++) {
            const int r = r0 + i;
            const int p = pp[i];
            if (p >= 0) {
                const double tt = ca[i] / sD[L(r)];
                sD[L(p)] -= cf[i] * tt;
                sB[L(p)] -= sB[L(r)] * tt;
            }
        }
    }
    for (int r1 = R - 1; r1 >= 0; r1 -= ACH) {
        int pp[ACH];
        double cb[ACH];
#pragma unroll
        for (int i = 0; i < ACH; i++) {
            const int r = r1 - i;
            pp[i] = -1;
            cb[i] = 0.0;
            if (r >= 0) {
                pp[i] = d.par[Q(r)];
                cb[i] = d.tcb[Q(r)];
            }
        }
#pragma unroll
        for (int i = 0; i < ACH; i++) {
            const int r = r1 - i;
            if (r >= 0) {
                const int p = pp[i];
                double bv = sB[L(r)];
                if (p >= 0) bv -= cb[i] * sX[L(p)];
                sX[L(r)] = bv / sD[L(r)];
            }
        }
    }
    // back-substitution in reverse: rows ascending (children before parents)
    for (int r0 = 0; r0 < R; r0 += ACH) {
        int pp[ACH];
        double cb[ACH], gb[ACH];
#pragma unroll
        for (int i = 0; i < ACH; i++) {
            const int r = r0 + i;
            pp[i] = -1;
            cb[i] = gb[i] = 0.0;
            if (r < R) {
                pp[i] = d.par[Q(r)];
                cb[i] = d.tcb[Q(r)];
                if (pp[i] >= 0) gb[i] = Gtb[Q(r)];
            }
        }
#pragma unroll
        for (int i = 0; i < ACH; i++) {
            const int r = r0 + i;
            if (r < R) {
                const int p = pp[i];
                const double xb = sY[L(r)];
                const double dk = sD[L(r)];
                const double bb = 0.0 + xb / dk;
                const double db = 0.0 - xb * sX[L(r)] / dk;
                if (p >= 0) {
                    Gtb[Q(r)] = gb[i] - xb * sX[L(p)] / dk;
                    sY[L(p)] -= xb * cb[i] / dk;
                }
                sY[L(r)] = bb;
                sX[L(r)] = db;
            }
        }
    }
    // forward elimination in reverse: rows descending
    for (int r1 = R - 2; r1 >= 0; r1 -= ACH) {
        int pp[ACH];
        double ca[ACH], cf[ACH], ga[ACH], gf[ACH];
#pragma unroll
        for (int i = 0; i < ACH; i++) {
            const int r = r1 - i;
            pp[i] = -1;
            ca[i] = cf[i] = ga[i] = gf[i] = 0.0;
            if (r >= 0) {
                pp[i] = d.par[Q(r)];
                if (pp[i] >= 0) {
                    ca[i] = d.ta[Q(r)];
                    cf[i] = d.tcf[Q(r)];
                    ga[i] = Gta[Q(r)];
                    gf[i] = Gtf[Q(r)];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < ACH; i++) {
            const int r = r1 - i;
            const int p = pp[i];
            if (p >= 0) {
                const double dk = sD[L(r)];
                const double tt = ca[i] / dk;
                const double bbp = sY[L(p)], dbp = sX[L(p)];
                sY[L(r)] -= bbp * tt;
                const double tbar = -bbp * sB[L(r)] - dbp * cf[i];
                Gtf[Q(r)] = gf[i] - dbp * tt;
                Gta[Q(r)] = ga[i] + tbar / dk;
                sX[L(r)] -= tbar * tt / dk;
            }
        }
    }
#pragma unroll 8
    for (int r = 0; r < R; r++) {
        a.BB[Q(r)] = sY[L(r)];
        a.DB[Q(r)] = sX[L(r)];
    }
#undef L
#undef Q
}

// The pack layout's reverse launch with the interpreter (hsh_adjoint_device.h).
__global__ __launch_bounds__(256) void k_adj_forest(AdjPack a)
{
    extern __shared__ double lds[];
    body_adj_forest<AdjInterp>(a, lds);
}

__global__ void k_gather(const double* __restrict__ V, const int* __restrict__ pos, int n, double* __restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = V[pos[j]];
}

__global__ void k_inject(double* __restrict__ LV, const int* __restrict__ pos, int n, const double* __restrict__ g)
{
    // one thread per observed row: rows are distinct, so no two threads
    // touch one position
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) LV[pos[j]] += g[j];
}

}  // namespace

hipError_t launch_adj_rhs(const AdjModel& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_rhs, dim3((a.d.nblocks + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adj_elim(const AdjModel& a, hipStream_t s)
{
    if (a.max_rows >= 1 && a.max_rows <= ADJ_LDS_ROWS) {
        const unsigned lds = (unsigned)(4 * a.max_rows * W * sizeof(double));
        hipLaunchKernelGGL(k_adj_elim_l, dim3(a.d.ntiles), dim3(64), lds, s, a);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_adj_elim, dim3((a.d.ntiles + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adj_local(const AdjModel& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_local, dim3((a.d.nblocks + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adj_forest(const AdjPack& a, const JitModule* jm, int max_pack_pos, hipStream_t s)
{
    if (a.d.npacks <= 0 || a.n <= 0) return hipSuccess;
    const unsigned lds = (unsigned)((size_t)ADJ_PACK_LDS_PER_POS * max_pack_pos + 16);
    if (jm && jm->ok && jm->adj_forest) {
        AdjPack aa = a;
        void* args[] = {&aa};
        return hipModuleLaunchKernel(jm->adj_forest, (unsigned)a.d.npacks, 1, 1, 256, 1, 1, lds, s, args, nullptr);
    }
    hipLaunchKernelGGL(k_adj_forest, dim3(a.d.npacks), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_gather(const double* V, const int* pos, int n, double* out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, V, pos, n, out);
    return hipGetLastError();
}

hipError_t launch_inject(double* LV, const int* pos, int n, const double* g, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_inject, dim3((n + 255) / 256), dim3(256), 0, s, LV, pos, n, g);
    return hipGetLastError();
}

}  // namespace hsh
