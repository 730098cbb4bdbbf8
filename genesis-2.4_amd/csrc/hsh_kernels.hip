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

// Slot access for run_prog.  With REG the compartment's whole slot vector is
// fetched with back-to-back coalesced loads before the program runs (one
// memory latency per compartment instead of one per micro-op) and indexed with
// the wave-uniform slot cursor, which gfx950 does in VGPRs (s_set_gpr_idx);
// gate states are written through to memory.  Without REG each slot is read
// when its micro-op runs (programs with more than MAXS slots).
constexpr int MAXS = 12;

// One compartment's channel program (device micro-ops, see hsh_internal.h).
// S and I point at this lane's slot 0; slot k lives at S[k*W].  code[0] is the
// program's slot count, the micro-ops follow.
template <bool REG>
__device__ __forceinline__ CompOut run_prog(const DevModel& d, const int32_t* __restrict__ code,
                                            double* __restrict__ Sb, int32_t* __restrict__ Ib,
                                            int lane, double Vm, bool pred_ok, double Vp,
                                            bool active, int step)
{
    // Sb / Ib: wave-uniform base of the block's slot area; slot c of this lane
    // is Sb[c*W + lane]
    double s[REG ? MAXS : 1];
    if (REG) {
        // unconditional loads (indices past the program's last slot re-read
        // that slot, which the cache serves), so they issue back to back
        const int last = code[0] - 1;
#pragma unroll
        for (int i = 0; i < MAXS; i++) {
            const int j = __builtin_amdgcn_readfirstlane(i < last ? i : last);
            s[i] = Sb[j * W + lane];
        }
    }
    int32_t* I = Ib + lane;
#define SGET(c) (REG ? s[(c)] : Sb[(c) * W + lane])
    // a slot is never read again after the program writes it (the cursor
    // only moves forward), so writes go to memory only and the register copy
    // stays read-only
#define SPUT(c, v) (Sb[(c) * W + lane] = (v))
    double sumgchan = 0.0;
    double ichan = SGET(0) + SGET(1);
    double Gk = 0.0, Ek = 0.0;
    double vipol = 0.0;
    const double* vtab = d.tab;
    const int ncols = d.ncols;
    int c = 2;
    int pc = 1;
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
            const double sv = SGET(c);
            if (k > 0) {
                X = (sv * (2.0 - X) + Y) / X;
            } else {
                X = Y / X;
                k = -k;
            }
            SPUT(c, X);
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
            Ek = SGET(c);
            Gk = SGET(c + 1);
            c += 2;
        } else if (op == U_ADDCURR) {
            sumgchan += Gk;
            ichan += Ek * Gk;
        } else if (op == U_NEWVOLT) {
            int filo;
            newvolt(d, Vm, pred_ok, Vp, filo, vipol);
            vtab = d.tab + (size_t)filo * ncols;
        } else if (op == U_CHAN) {
            Gk = SGET(c);
            c += 1;
        } else if (op == U_SPIKE) {
            const int is = code[pc++];
            int cnt = I[is * W] - 1;
            if (Vm > SGET(c)) {
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
            const double tbyc = SGET(c);
            CompOut o;
            o.rhs = Vm + ichan * tbyc;
            o.diag = sumgchan * tbyc + SGET(c + 1);
            return o;
        }
    }
#undef SGET
#undef SPUT
}



// ---------------------------------------------------------------------------
// Plain drop-in, part (a): one wave per block, all blocks in one launch.
template <bool REG>
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
    const int step = d.step[bm.tile];
    CompOut o = run_prog<REG>(d, d.prog + bm.prog, d.S + bm.dslot, d.I + bm.islot, lane, Vm,
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
    if (t >= d.ntiles) return;
    const int lane = threadIdx.x & (W - 1);
    if (lane == 0) d.step[t] += 1;
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

template <bool REG>
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
    const int step0 = d.step[t];
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
            CompOut o = run_prog<REG>(d, d.prog + bm.prog, d.S + bm.dslot, d.I + bm.islot, lane,
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
    if (threadIdx.x == 0) d.step[t] = step0 + K;
}

// Batched persistent kernel for small trees (rows <= 4*RPW): four waves per
// tile, wave w owns rows w, w+4, ...  Each owner keeps its rows' tree
// coefficients in registers for the whole launch, and the elimination walks
// the rows in order -- the owner of row r folds it into its parent, then all
// four waves meet at a barrier -- so the tile needs only 24 bytes of LDS per
// compartment and every tile of the workload is resident at once.
template <int RPW>
__global__ __launch_bounds__(256) void k_batched_r(DevModel d, int K)
{
    constexpr int NW = 4;
    constexpr int RMAX = NW * RPW;
    extern __shared__ double lds[];
    const int t = blockIdx.x;
    const TileMeta tm = d.tiles[t];
    const int R = tm.rows, b0 = tm.block0;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & (W - 1);
    const bool active = lane < tm.lanes;
    double* LV = lds;
    double* LB = lds + R * W;
    double* LD = lds + 2 * R * W;
    int par[RPW], pidx[RPW];
    double ta[RPW], tcf[RPW], tcb[RPW];
#pragma unroll
    for (int i = 0; i < RPW; i++) {
        const int r = wave + NW * i;
        par[i] = -1; pidx[i] = -1;
        ta[i] = tcf[i] = tcb[i] = 0.0;
        if (r < R) {
            const size_t q = (size_t)(b0 + r) * W + lane;
            par[i] = d.par[q];
            ta[i] = d.ta[q];
            tcf[i] = d.tcf[q];
            tcb[i] = d.tcb[q];
            LV[r * W + lane] = d.V[q];
            const int pp = d.pred[q];
            if (pp >= 0 && (pp >> 6) >= b0 && (pp >> 6) < b0 + R)
                pidx[i] = ((pp >> 6) - b0) * W + (pp & (W - 1));
        }
    }
    const bool is_last = ((size_t)(b0 + R - 1) * W + lane) == (size_t)d.last_pos;
    const int step0 = d.step[t];
    __syncthreads();
    for (int k = 0; k < K; k++) {
#pragma unroll 1
        for (int i = 0; i < RPW; i++) {
            const int r = wave + NW * i;
            if (r < R) {
                const int b = b0 + r;
                const BlockMeta bm = d.blocks[b];
                const double Vm = LV[r * W + lane];
                const bool pok = pidx[i] >= 0;
                const double Vp = pok ? LV[pidx[i]] : 0.0;
                CompOut o = run_prog<true>(d, d.prog + bm.prog, d.S + bm.dslot, d.I + bm.islot, lane,
                                     Vm, pok, Vp, active, step0 + k);
                LB[r * W + lane] = o.rhs;
                LD[r * W + lane] = o.diag;
            }
        }
        __syncthreads();
        double droot0 = 0.0;
        if (wave == ((R - 1) & (NW - 1))) droot0 = LD[(R - 1) * W + lane];
#pragma unroll
        for (int r = 0; r < RMAX - 1; r++) {
            if (r < R - 1) {
            if (wave == (r & (NW - 1))) {
                const int i = r / NW;
                const int p = par[i];
                if (p >= 0) {
                    const double tt = ta[i] / LD[r * W + lane];
                    LD[p * W + lane] -= tcf[i] * tt;
                    LB[p * W + lane] -= LB[r * W + lane] * tt;
                }
            }
            __syncthreads();
            }
        }
#pragma unroll
        for (int r = RMAX - 1; r >= 0; r--) {
            if (r < R) {
                if (wave == (r & (NW - 1))) {
                    const int i = r / NW;
                    const int p = par[i];
                    double bv = LB[r * W + lane];
                    if (p >= 0) bv -= tcb[i] * LB[p * W + lane];
                    const double x = bv / LD[r * W + lane];
                    LB[r * W + lane] = x;
                    if (d.crank) {
                        const double v = LV[r * W + lane];
                        LV[r * W + lane] = x + x - v;
                    } else {
                        LV[r * W + lane] = x;
                    }
                    if (r == R - 1 && is_last) LD[r * W + lane] = droot0;
                }
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int i = 0; i < RPW; i++) {
        const int r = wave + NW * i;
        if (r < R) {
            const size_t b = (size_t)(b0 + r);
            d.V[b * W + lane] = LV[r * W + lane];
            d.R[b * 2 * W + lane] = LB[r * W + lane];
            d.R[b * 2 * W + W + lane] = LD[r * W + lane];
        }
    }
    if (threadIdx.x == 0) d.step[t] = step0 + K;
}

// Plain drop-in elimination for trees of up to RMAX rows: one wave per tile,
// the tile's rhs/diag staged in LDS (each lane touches only its own column),
// tree coefficients loaded into registers before the sweep starts.
template <int RMAX>
__global__ __launch_bounds__(64) void k_elim_r(DevModel d)
{
    __shared__ double sB[RMAX * W], sD[RMAX * W];
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    const TileMeta tm = d.tiles[t];
    const int R = tm.rows, b0 = tm.block0;
    if (lane == 0) d.step[t] += 1;
    double* Rr = d.R + (size_t)b0 * 2 * W + lane;
    const size_t p0 = (size_t)b0 * W + lane;
    int par[RMAX];
    double ta[RMAX], tcf[RMAX], tcb[RMAX];
#pragma unroll
    for (int r = 0; r < RMAX; r++) {
        if (r < R) {
            const size_t q = p0 + (size_t)r * W;
            par[r] = d.par[q];
            ta[r] = d.ta[q];
            tcf[r] = d.tcf[q];
            tcb[r] = d.tcb[q];
            sB[r * W + lane] = Rr[(size_t)r * 2 * W];
            sD[r * W + lane] = Rr[(size_t)r * 2 * W + W];
        } else {
            par[r] = -1;
            ta[r] = tcf[r] = tcb[r] = 0.0;
        }
    }
    const bool is_last = (p0 + (size_t)(R - 1) * W) == (size_t)d.last_pos;
    const double droot0 = sD[(R - 1) * W + lane];
#pragma unroll
    for (int r = 0; r < RMAX - 1; r++) {
        if (r < R - 1) {
            const int p = par[r];
            if (p >= 0) {
                const double tt = ta[r] / sD[r * W + lane];
                sD[p * W + lane] -= tcf[r] * tt;
                sB[p * W + lane] -= sB[r * W + lane] * tt;
            }
        }
    }
#pragma unroll
    for (int r = RMAX - 1; r >= 0; r--) {
        if (r < R) {
            const size_t q = p0 + (size_t)r * W;
            double bv = sB[r * W + lane];
            const int p = par[r];
            if (p >= 0) bv -= tcb[r] * sB[p * W + lane];
            const double x = bv / sD[r * W + lane];
            sB[r * W + lane] = x;
            if (d.crank) {
                const double v = d.V[q];
                d.V[q] = x + x - v;
            } else {
                d.V[q] = x;
            }
        }
    }
    if (is_last) sD[(R - 1) * W + lane] = droot0;
#pragma unroll
    for (int r = 0; r < RMAX; r++) {
        if (r < R) {
            Rr[(size_t)r * 2 * W] = sB[r * W + lane];
            Rr[(size_t)r * 2 * W + W] = sD[r * W + lane];
        }
    }
}

// Isopotential cells (every tree one compartment): one lane per cell, the
// voltage stays in a register for K steps.  V = rhs/diag is the reference's
// single-row solve (hines_solve.c:236-238; CN :158-162).
template <bool REG>
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
    const int step0 = d.step[t];
    CompOut o{0.0, 1.0};
    double x = 0.0;
    for (int k = 0; k < K; k++) {
        const double Vp = __shfl(V, plane, W);
        o = run_prog<REG>(d, d.prog + bm.prog, d.S + bm.dslot, d.I + bm.islot, lane, V, pok, Vp,
                     lane < tm.lanes, step0 + k);
        x = o.rhs / o.diag;
        if (d.crank) V = x + x - V;
        else V = x;
    }
    d.V[pos] = V;
    // a one-row solver keeps rhs in results[0] (hines_solve.c:236-238)
    d.R[(size_t)b * 2 * W + lane] = d.single ? o.rhs : x;
    d.R[(size_t)b * 2 * W + W + lane] = o.diag;
    if (lane == 0) d.step[t] = step0 + K;
}

__global__ void k_step_advance(int32_t* step, int32_t n, int32_t by)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) step[t] += by;
}

}  // namespace

int batched_lds_bytes(int max_rows) { return 3 * max_rows * W * (int)sizeof(double); }

hipError_t launch_channel(const DevModel& d, hipStream_t s)
{
    dim3 grid((d.nblocks + 3) / 4);
    if (d.max_slots <= MAXS)
        hipLaunchKernelGGL(k_channel<true>, grid, dim3(256), 0, s, d);
    else
        hipLaunchKernelGGL(k_channel<false>, grid, dim3(256), 0, s, d);
    return hipGetLastError();
}

hipError_t launch_elim(const DevModel& d, int max_rows, hipStream_t s)
{
    if (max_rows <= 16) {
        hipLaunchKernelGGL(k_elim_r<16>, dim3(d.ntiles), dim3(64), 0, s, d);
    } else {
        dim3 grid((d.ntiles + 3) / 4);
        hipLaunchKernelGGL(k_elim, grid, dim3(256), 0, s, d);
    }
    return hipGetLastError();
}

hipError_t launch_batched(const DevModel& d, int ksteps, int waves, int max_rows, hipStream_t s)
{
    if (max_rows > BATCH_MAX_ROWS) return hipErrorInvalidValue;
    const bool reg = d.max_slots <= MAXS;
    if (waves == 0 && max_rows <= 16 && reg) {
        const int lds = batched_lds_bytes(max_rows);
        if (max_rows <= 4)
            hipLaunchKernelGGL(k_batched_r<1>, dim3(d.ntiles), dim3(256), lds, s, d, ksteps);
        else if (max_rows <= 8)
            hipLaunchKernelGGL(k_batched_r<2>, dim3(d.ntiles), dim3(256), lds, s, d, ksteps);
        else
            hipLaunchKernelGGL(k_batched_r<4>, dim3(d.ntiles), dim3(256), lds, s, d, ksteps);
        return hipGetLastError();
    }
    if (waves < 1) waves = 1;
    if (waves > 16) waves = 16;
    if (reg)
        hipLaunchKernelGGL(k_batched<true>, dim3(d.ntiles), dim3(waves * W), batched_lds_bytes(max_rows),
                           s, d, ksteps);
    else
        hipLaunchKernelGGL(k_batched<false>, dim3(d.ntiles), dim3(waves * W), batched_lds_bytes(max_rows),
                           s, d, ksteps);
    return hipGetLastError();
}

hipError_t launch_iso(const DevModel& d, int ksteps, hipStream_t s)
{
    dim3 grid((d.ntiles + 3) / 4);
    if (d.max_slots <= MAXS)
        hipLaunchKernelGGL(k_iso<true>, grid, dim3(256), 0, s, d, ksteps);
    else
        hipLaunchKernelGGL(k_iso<false>, grid, dim3(256), 0, s, d, ksteps);
    return hipGetLastError();
}

hipError_t launch_step_advance(const DevModel& d, int32_t by, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_advance, dim3((d.ntiles + 255) / 256), dim3(256), 0, s, d.step, d.ntiles, by);
    return hipGetLastError();
}

}  // namespace hsh
