// hsh_kernels.hip -- gfx950 kernels of libhsolve_hip.
//
// Data layout (HBM): every array is [block][slot][lane] with 64 lanes per
// block, a block being one row of a tile of 64 neurons that share size and
// per-row channel programs.  A wave owns a block (channel kernels) or a tile
// (elimination), so all loads of one slot by one wave are one contiguous
// 512-byte line and the program a wave interprets is wave-uniform.
//
// Arithmetic follows the reference CPU code expression by expression
// (genesis/src/hines/hines_chip.c:157-600, hines_solve.c:140-296) and the
// library is compiled with -ffp-contract=off, so each result is rounded as the
// reference rounds it; fp64 division is IEEE correctly rounded.
#include "hsh_internal.h"

namespace hsh {

namespace {

struct CompOut {
    double rhs, diag;
};

// Table row and fraction for voltage Vm (hines_chip.c:174-203).  The
// reference keeps the row of the previous NEWVOLT compartment (xlo/xhi cache)
// and reuses it when Vm falls inside that row's [xvals[f], xvals[f+1]);
// the previous compartment's own row is recomputed here from its voltage.
__device__ __forceinline__ void newvolt(const DevModel& d, double Vm, bool pred_ok, double Vp,
                                        int& filo, double& vipol)
{
    double vraw = (Vm - d.xmin) * d.invdx;
    if (pred_ok) {
        double pr = (Vp - d.xmin) * d.invdx;
        if (pr < 0.0) pr = 0.0;
        else if (pr > d.xdivs) pr = d.xdivs;
        int fp = (int)pr;
        fp = fp < 0 ? 0 : (fp > d.xdivs ? d.xdivs : fp);
        double xlo = d.xvals[fp], xhi = d.xvals[fp + 1];
        if (!((xlo > Vm) || (Vm >= xhi))) {
            filo = fp;
            vipol = vraw - (double)fp;
            return;
        }
    }
    double v = vraw;
    if (v < 0.0) v = 0.0;
    else if (v > d.xdivs) v = d.xdivs;
    int f = (int)v;
    f = f < 0 ? 0 : (f > d.xdivs ? d.xdivs : f);
    filo = f;
    vipol = v - (double)f;
}

// One compartment's channel program (device micro-ops, see hsh_internal.h).
// S and I point at this lane's slot 0; slot k lives at S[k*W].
__device__ __forceinline__ CompOut run_comp(const DevModel& d, const int32_t* __restrict__ code,
                                            double* __restrict__ S, int32_t* __restrict__ I,
                                            double Vm, bool pred_ok, double Vp, bool active,
                                            int step)
{
    double sumgchan = 0.0;
    double ichan = S[0] + S[W];
    double Gk = 0.0, Ek = 0.0;
    double vipol = 0.0;
    const double* vtab = d.tab;
    const int ncols = d.ncols;
    int c = 2;
    int pc = 0;
    for (;;) {
        const int op = code[pc++];
        if (op == U_IPOL1V) {
            const int col = code[pc];
            int k = code[pc + 1];
            pc += 2;
            const double* tab = vtab + col;
            double X = tab[0];
            X = X + vipol * (tab[ncols] - X);
            double Y = tab[1];
            Y = Y + vipol * (tab[ncols + 1] - Y);
            double s = S[c * W];
            if (k > 0) {
                X = (s * (2.0 - X) + Y) / X;
            } else {
                X = Y / X;
                k = -k;
            }
            S[c * W] = X;
            c++;
            if (k == 1) {
                Gk *= X;
            } else if (k == 2) {
                Gk *= X * X;
            } else if (k == 3) {
                Gk *= X * X * X;
            } else if (k == 4) {
                X *= X;
                Gk *= X * X;
            } else if (k == 5) {
                Gk *= X;
                X *= X;
                X *= X;
                Gk *= X;
            } else {
                X *= X;
                Gk *= X;
                X *= X;
                Gk *= X;
            }
        } else if (op == U_CHAN_EK) {
            Ek = S[c * W];
            Gk = S[(c + 1) * W];
            c += 2;
        } else if (op == U_ADDCURR) {
            sumgchan += Gk;
            ichan += Ek * Gk;
        } else if (op == U_NEWVOLT) {
            int filo;
            newvolt(d, Vm, pred_ok, Vp, filo, vipol);
            vtab = d.tab + (size_t)filo * ncols;
        } else if (op == U_CHAN) {
            Gk = S[c * W];
            c += 1;
        } else if (op == U_SPIKE) {
            const int is = code[pc++];
            int cnt = I[is * W] - 1;
            if (Vm > S[c * W]) {
                if (cnt <= 0) {
                    if (active) {
                        int q = atomicAdd(d.nspk, 1);
                        if (q < d.spk_cap) {
                            d.spk[2 * q] = I[(is + 2) * W];
                            d.spk[2 * q + 1] = step;
                        }
                    }
                    cnt = I[(is + 1) * W];
                }
            }
            c += 1;
            I[is * W] = cnt;
        } else {  // U_END
            const double tbyc = S[c * W];
            CompOut o;
            o.rhs = Vm + ichan * tbyc;
            o.diag = sumgchan * tbyc + S[(c + 1) * W];
            return o;
        }
    }
}

// ---------------------------------------------------------------------------
// Plain drop-in, part (a): one wave per block, all blocks in one launch.
__global__ __launch_bounds__(256) void k_channel(DevModel d)
{
    int b = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    b = __builtin_amdgcn_readfirstlane(b);
    if (b >= d.nblocks) return;
    const int lane = threadIdx.x & (W - 1);
    const BlockMeta bm = d.blocks[b];
    const TileMeta tm = d.tiles[bm.tile];
    const int pos = b * W + lane;
    const double Vm = d.V[pos];
    const int pp = d.pred[pos];
    const double Vp = pp >= 0 ? d.V[pp] : 0.0;
    const int step = d.step[0];
    CompOut o = run_comp(d, d.prog + bm.prog, d.S + bm.dslot + lane, d.I + bm.islot + lane, Vm,
                         pp >= 0, Vp, lane < tm.lanes, step);
    d.R[(size_t)b * 2 * W + lane] = o.rhs;
    d.R[(size_t)b * 2 * W + W + lane] = o.diag;
}

// Plain drop-in, part (b): one lane per neuron, one wave per tile.  Forward
// elimination in push form -- when row r is final it is folded into its
// parent -- which applies each parent's kids in ascending row order, the
// order of the reference's funcs[] program; then back-substitution.
__global__ __launch_bounds__(256) void k_elim(DevModel d)
{
    int t = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    t = __builtin_amdgcn_readfirstlane(t);
    if (blockIdx.x == 0 && threadIdx.x == 0) d.step[0] += 1;
    if (t >= d.ntiles) return;
    const int lane = threadIdx.x & (W - 1);
    const TileMeta tm = d.tiles[t];
    const int R = tm.rows, b0 = tm.block0;
    double* Rr = d.R + (size_t)b0 * 2 * W + lane;     // row r: rhs Rr[r*2W], diag Rr[r*2W+W]
    const size_t p0 = (size_t)b0 * W + lane;
    // The reference never stores the eliminated diagonal of its last row
    // (hines_solve.c:275-276 writes only the solution), so that slot keeps
    // the channel update's value.
    const bool is_last = (p0 + (size_t)(R - 1) * W) == (size_t)d.last_pos;
    const double droot0 = Rr[(size_t)(R - 1) * 2 * W + W];
    for (int r = 0; r < R - 1; r++) {
        const int p = d.par[p0 + (size_t)r * W];
        if (p >= 0) {
            const double dk = Rr[(size_t)r * 2 * W + W];
            const double bk = Rr[(size_t)r * 2 * W];
            const double tt = d.ta[p0 + (size_t)r * W] / dk;
            Rr[(size_t)p * 2 * W + W] -= d.tcf[p0 + (size_t)r * W] * tt;
            Rr[(size_t)p * 2 * W] -= bk * tt;
        }
    }
    for (int r = R - 1; r >= 0; r--) {
        const size_t q = p0 + (size_t)r * W;
        double bv = Rr[(size_t)r * 2 * W];
        const int p = d.par[q];
        if (p >= 0) bv -= d.tcb[q] * Rr[(size_t)p * 2 * W];
        const double x = bv / Rr[(size_t)r * 2 * W + W];
        Rr[(size_t)r * 2 * W] = x;
        if (d.crank) {
            const double v = d.V[q];
            d.V[q] = x + x - v;
        } else {
            d.V[q] = x;
        }
    }
    if (is_last) Rr[(size_t)(R - 1) * 2 * W + W] = droot0;
}

// ---------------------------------------------------------------------------
// Batched persistent kernel: one workgroup per tile, K steps per launch.
// Voltages, right-hand sides and diagonals of the tile stay in LDS for the
// whole launch; waves split the rows for the channel phase, wave 0 runs the
// 64 eliminations (one per lane) between two barriers.
constexpr int BATCH_MAX_ROWS = 64;

__global__ __launch_bounds__(1024) void k_batched(DevModel d, int K)
{
    extern __shared__ double lds[];
    const int t = blockIdx.x;
    const TileMeta tm = d.tiles[t];
    const int R = tm.rows, b0 = tm.block0;
    const int nw = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & (W - 1);
    double* LV = lds;
    double* LB = lds + R * W;
    double* LD = lds + 2 * R * W;
    for (int r = wave; r < R; r += nw) LV[r * W + lane] = d.V[(size_t)(b0 + r) * W + lane];
    const int step0 = d.step[0];
    __syncthreads();
    for (int k = 0; k < K; k++) {
        for (int r = wave; r < R; r += nw) {
            const int b = b0 + r;
            const BlockMeta bm = d.blocks[b];
            const int pos = b * W + lane;
            const double Vm = LV[r * W + lane];
            const int pp = d.pred[pos];
            bool pok = false;
            double Vp = 0.0;
            if (pp >= 0) {
                const int pb = pp >> 6;
                if (pb >= b0 && pb < b0 + R) {
                    pok = true;
                    Vp = LV[(pb - b0) * W + (pp & (W - 1))];
                }
            }
            CompOut o = run_comp(d, d.prog + bm.prog, d.S + bm.dslot + lane, d.I + bm.islot + lane,
                                 Vm, pok, Vp, lane < tm.lanes, step0 + k);
            LB[r * W + lane] = o.rhs;
            LD[r * W + lane] = o.diag;
        }
        __syncthreads();
        if (wave == 0) {
            const size_t p0 = (size_t)b0 * W + lane;
            const bool is_last = (p0 + (size_t)(R - 1) * W) == (size_t)d.last_pos;
            const double droot0 = LD[(R - 1) * W + lane];
            for (int r = 0; r < R - 1; r++) {
                const int p = d.par[p0 + (size_t)r * W];
                if (p >= 0) {
                    const double tt = d.ta[p0 + (size_t)r * W] / LD[r * W + lane];
                    LD[p * W + lane] -= d.tcf[p0 + (size_t)r * W] * tt;
                    LB[p * W + lane] -= LB[r * W + lane] * tt;
                }
            }
            for (int r = R - 1; r >= 0; r--) {
                const size_t q = p0 + (size_t)r * W;
                double bv = LB[r * W + lane];
                const int p = d.par[q];
                if (p >= 0) bv -= d.tcb[q] * LB[p * W + lane];
                const double x = bv / LD[r * W + lane];
                LB[r * W + lane] = x;
                if (d.crank) {
                    const double v = LV[r * W + lane];
                    LV[r * W + lane] = x + x - v;
                } else {
                    LV[r * W + lane] = x;
                }
            }
            if (is_last) LD[(R - 1) * W + lane] = droot0;
        }
        __syncthreads();
    }
    for (int r = wave; r < R; r += nw) {
        const size_t b = (size_t)(b0 + r);
        d.V[b * W + lane] = LV[r * W + lane];
        d.R[b * 2 * W + lane] = LB[r * W + lane];
        d.R[b * 2 * W + W + lane] = LD[r * W + lane];
    }
}

// Isopotential cells (every tree one compartment): one lane per cell, the
// voltage stays in a register for K steps.  V = rhs/diag is the reference's
// single-row solve (hines_solve.c:236-238; CN :158-162).
__global__ __launch_bounds__(256) void k_iso(DevModel d, int K)
{
    int t = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= d.ntiles) return;
    const int lane = threadIdx.x & (W - 1);
    const TileMeta tm = d.tiles[t];
    const int b = tm.block0;
    const BlockMeta bm = d.blocks[b];
    const int pos = b * W + lane;
    double V = d.V[pos];
    const int pp = d.pred[pos];
    // the previous cell in op order, if it sits in this wave
    const bool pok = pp >= 0 && (pp >> 6) == b;
    const int plane = pok ? (pp & (W - 1)) : lane;
    const int step0 = d.step[0];
    CompOut o{0.0, 1.0};
    double x = 0.0;
    for (int k = 0; k < K; k++) {
        const double Vp = __shfl(V, plane, W);
        o = run_comp(d, d.prog + bm.prog, d.S + bm.dslot + lane, d.I + bm.islot + lane, V, pok, Vp,
                     lane < tm.lanes, step0 + k);
        x = o.rhs / o.diag;
        if (d.crank) V = x + x - V;
        else V = x;
    }
    d.V[pos] = V;
    // a one-row solver keeps rhs in results[0] (hines_solve.c:236-238)
    d.R[(size_t)b * 2 * W + lane] = d.single ? o.rhs : x;
    d.R[(size_t)b * 2 * W + W + lane] = o.diag;
}

__global__ void k_step_advance(int32_t* step, int32_t by)
{
    if (threadIdx.x == 0) step[0] += by;
}

}  // namespace

int batched_lds_bytes(int max_rows) { return 3 * max_rows * W * (int)sizeof(double); }

hipError_t launch_channel(const DevModel& d, hipStream_t s)
{
    dim3 grid((d.nblocks + 3) / 4);
    hipLaunchKernelGGL(k_channel, grid, dim3(256), 0, s, d);
    return hipGetLastError();
}

hipError_t launch_elim(const DevModel& d, hipStream_t s)
{
    dim3 grid((d.ntiles + 3) / 4);
    hipLaunchKernelGGL(k_elim, grid, dim3(256), 0, s, d);
    return hipGetLastError();
}

hipError_t launch_batched(const DevModel& d, int ksteps, int waves, int max_rows, hipStream_t s)
{
    if (max_rows > BATCH_MAX_ROWS) return hipErrorInvalidValue;
    if (waves < 1) waves = 1;
    if (waves > 16) waves = 16;
    hipLaunchKernelGGL(k_batched, dim3(d.ntiles), dim3(waves * W), batched_lds_bytes(max_rows), s, d,
                       ksteps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, s, d.step, ksteps);
    return hipGetLastError();
}

hipError_t launch_iso(const DevModel& d, int ksteps, hipStream_t s)
{
    dim3 grid((d.ntiles + 3) / 4);
    hipLaunchKernelGGL(k_iso, grid, dim3(256), 0, s, d, ksteps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, s, d.step, ksteps);
    return hipGetLastError();
}

hipError_t launch_step_advance(int32_t* step, int32_t by, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, s, step, by);
    return hipGetLastError();
}

}  // namespace hsh
