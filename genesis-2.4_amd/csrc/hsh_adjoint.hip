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

namespace hsh {

namespace {

constexpr int MAXG = 3;     // gates per channel the reverse program keeps
constexpr int ADJ_LDS_ROWS = 64;   // k_adj_elim_l: 4 x 64 x 64 doubles = 128 KiB of LDS

struct Interp {
    int filo;
    double vipol;
    bool inside;
};

__device__ __forceinline__ Interp table_row(const DevModel& d, double Vm)
{
    Interp r;
    const double raw = (Vm - d.xmin) * d.invdx;
    double v = raw;
    if (v < 0.0) v = 0.0;
    else if (v > d.xdivs) v = d.xdivs;
    int f = (int)v;
    f = f < 0 ? 0 : (f > d.xdivs ? d.xdivs : f);
    r.filo = f;
    r.vipol = v - (double)f;
    r.inside = raw > 0.0 && raw < (double)d.xdivs;
    return r;
}

// rhs and diag of one compartment (the forward program, no slot writes).
__device__ CompOut adj_forward_local(const DevModel& d, const int* __restrict__ code, const double* __restrict__ Sb,
                                     int lane, double Vm)
{
    double sumg = 0.0;
    double ichan = Sb[0 * W + lane] + Sb[1 * W + lane];
    double Gk = 0.0, Ek = 0.0;
    Interp ip{0, 0.0, false};
    const double* vtab = d.tab;
    const int ncols = d.ncols;
    int c = 2, pc = PROG_HEADER;
    for (;;) {
        const int op = code[pc++];
        if (op == U_IPOL1V) {
            const int col = code[pc], k = code[pc + 1];
            pc += 2;
            const double X = gate_update(vtab, col, ncols, ip.vipol, Sb[c * W + lane], k < 0);
            c++;
            Gk = apply_power(Gk, X, k < 0 ? -k : k);
        } else if (op == U_CHAN_EK) {
            Ek = Sb[c * W + lane];
            Gk = Sb[(c + 1) * W + lane];
            c += 2;
        } else if (op == U_ADDCURR) {
            sumg += Gk;
            ichan += Ek * Gk;
        } else if (op == U_NEWVOLT) {
            ip = table_row(d, Vm);
            vtab = d.tab + (long long)ip.filo * ncols;
        } else if (op == U_CHAN) {
            Gk = Sb[c * W + lane];
            c += 1;
        } else if (op == U_SPIKE) {
            pc++;
            c += 1;
        } else {
            const double tbyc = Sb[c * W + lane];
            CompOut o;
            o.rhs = Vm + ichan * tbyc;
            o.diag = sumg * tbyc + Sb[(c + 1) * W + lane];
            return o;
        }
    }
}

__device__ __forceinline__ double ipow(double x, int p)
{
    double r = 1.0;
    for (int i = 0; i < p; i++) r *= x;
    return r;
}

// Reverse of one compartment's program.  bb, db: adjoints of its rhs and
// diag; LSb: the slot adjoints (gate slots hold the adjoint of the gate state
// at t+1 on entry and at t on return; other slots accumulate gradients).
// Returns the adjoint of Vm at t contributed through rhs and the gates.
__device__ double adj_reverse_local(const DevModel& d, const int* __restrict__ code, const double* __restrict__ Sb,
                                    double* __restrict__ LSb, int lane, double Vm, double bb, double db)
{
    const int ns = code[0];
    const double tbyc = Sb[(ns - 2) * W + lane];
    const double lam_i = bb * tbyc;      // adjoint of ichan
    const double lam_g = db * tbyc;      // adjoint of sumgchan
    double vbar = bb;                    // rhs = Vm + ...
    double sumg = 0.0;
    double ichan = Sb[0 * W + lane] + Sb[1 * W + lane];
    LSb[0 * W + lane] += lam_i;          // Em/Rm
    LSb[1 * W + lane] += lam_i;          // inject
    double Gk = 0.0, Ek = 0.0, gmax = 0.0;
    int ekslot = -1, gslot = -1, ng = 0;
    int gs[MAXG], gp[MAXG];
    double gx[MAXG], gdx[MAXG], gdv[MAXG];
    Interp ip{0, 0.0, false};
    const double* vtab = d.tab;
    const int ncols = d.ncols;
    int c = 2, pc = PROG_HEADER;
    for (;;) {
        const int op = code[pc++];
        if (op == U_IPOL1V) {
            const int col = code[pc], k = code[pc + 1];
            pc += 2;
            const double* tab = vtab + col;
            const double B = tab[0] + ip.vipol * (tab[ncols] - tab[0]);
            const double dB = ip.inside ? (tab[ncols] - tab[0]) * d.invdx : 0.0;
            const double dA = ip.inside ? (tab[ncols + 1] - tab[1]) * d.invdx : 0.0;
            const double X0 = Sb[c * W + lane];
            double X, dXdX, dXdV;
            if (k < 0) {
                X = gate_update(vtab, col, ncols, ip.vipol, X0, true);
                dXdX = 0.0;
                dXdV = (dA - X * dB) / B;
            } else {
                X = gate_update(vtab, col, ncols, ip.vipol, X0, false);
                dXdX = (2.0 - B) / B;
                dXdV = (dA - X0 * dB - X * dB) / B;
            }
            const int p = k < 0 ? -k : k;
            Gk = apply_power(Gk, X, p);
            if (ng < MAXG) {
                gs[ng] = c;
                gp[ng] = p;
                gx[ng] = X;
                gdx[ng] = dXdX;
                gdv[ng] = dXdV;
            }
            ng++;
            c++;
        } else if (op == U_CHAN_EK) {
            ekslot = c;
            Ek = Sb[c * W + lane];
            gslot = c + 1;
            gmax = Gk = Sb[(c + 1) * W + lane];
            c += 2;
            ng = 0;
        } else if (op == U_CHAN) {
            gslot = c;
            gmax = Gk = Sb[c * W + lane];
            c += 1;
            ng = 0;
        } else if (op == U_ADDCURR) {
            sumg += Gk;
            ichan += Ek * Gk;
            const double lamG = lam_i * Ek + lam_g;
            if (ekslot >= 0) LSb[ekslot * W + lane] += lam_i * Gk;
            double prod = 1.0;
            for (int g = 0; g < ng && g < MAXG; g++) prod *= ipow(gx[g], gp[g]);
            LSb[gslot * W + lane] += lamG * prod;
            for (int g = 0; g < ng && g < MAXG; g++) {
                double dG = gmax * (double)gp[g] * ipow(gx[g], gp[g] - 1);
                for (int h = 0; h < ng && h < MAXG; h++)
                    if (h != g) dG *= ipow(gx[h], gp[h]);
                const double lamX = LSb[gs[g] * W + lane] + lamG * dG;
                LSb[gs[g] * W + lane] = lamX * gdx[g];
                vbar += lamX * gdv[g];
            }
            ng = 0;
        } else if (op == U_NEWVOLT) {
            ip = table_row(d, Vm);
            vtab = d.tab + (long long)ip.filo * ncols;
        } else if (op == U_SPIKE) {
            pc++;
            c += 1;
        } else {
            LSb[c * W + lane] += bb * ichan + db * sumg;     // tbyc
            LSb[(c + 1) * W + lane] += db;                   // diagterm
            return vbar;
        }
    }
}

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

// Every reverse step of one checkpoint segment on the pack layout, one
// workgroup per pack (neurons are independent blocks of the system, so a
// pack's adjoint never reads another pack).  The pack's working arrays, its
// level schedule and its tree coefficients are in LDS for the whole launch
// (ADJ_PACK_LDS_PER_POS bytes per position).  Per step t, last first:
//   * dL/dV of the observed rows enters the adjoint of V^{t+1};
//   * rhs/diag of step t from the recorded V^t, S^t (adj_forward_local);
//   * the elimination as the step ran it: forward by height (a parent folds
//     its kids in ascending row order), back-substitution by depth;
//   * the back-substitution in reverse, by height: a parent gathers its
//     kids' terms xb*tcb/dk in ascending row order -- the order in which
//     k_adj_elim's ascending rows subtract them -- and then every row's
//     adjoints of rhs and diag and the tcb gradient follow from its final
//     value;
//   * the forward elimination in reverse, by depth: a row reads its parent's
//     final adjoints of rhs and diag;
//   * each compartment's program in reverse (adj_reverse_local).
// Every value is computed by the operations of k_adj_elim in the same
// order, so the gradients equal the tile layout's bit for bit.  A pack whose
// levels are all at most one wave wide runs the sweeps in wave 0 alone (one
// wave's LDS accesses complete in order: no workgroup barrier per level);
// otherwise every wave takes part, with a barrier after each level.
__global__ __launch_bounds__(256) void k_adj_forest(AdjPack a)
{
    extern __shared__ double lds[];
    const DevModel& d = a.d;
    const PackMeta pm = d.packs[blockIdx.x];
    const int NP = pm.nblk * W;
    const int T = blockDim.x, tid = threadIdx.x;
    const int nw = T >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & (W - 1);
    const long long base = (long long)pm.block0 * W;
    double* sV = lds;            // adjoint of V
    double* sB = sV + NP;        // rhs, then the eliminated rhs
    double* sD = sB + NP;        // diag, then the eliminated diag
    double* sX = sD + NP;        // solution
    double* sY = sX + NP;        // adjoint of the solution, then of rhs
    double* sZ = sY + NP;        // adjoint of diag
    double* gA = sZ + NP;        // gradients of ta, tcf, tcb
    double* gF = gA + NP;
    double* gB = gF + NP;
    double* cA = gB + NP;        // the row's ta, tcf, tcb
    double* cF = cA + NP;
    double* cB = cF + NP;
    int* sPar = (int*)(cB + NP); // pack-local parent, -1 root, -2 padding
    int* sObs = sPar + NP;       // observed row index, -1
    int* sLvl = sObs + NP;       // level bounds: forward, then backward
    int* sFE = sLvl + (pm.nfl + pm.nbl + 2);
    int* sBK = sFE + 2 * pm.nfe;
    unsigned short* sKI = (unsigned short*)(sBK + pm.nbe);
    {
        const int* lvl = d.lvl + pm.lv0;
        const int* FE = d.fent + 2 * (long long)pm.fe0;
        const unsigned short* KI = d.kidx + pm.ke0;
        const int* BKP = d.bkp + pm.be0;
        for (int i = tid; i < NP; i += T) {
            sV[i] = a.LV[base + i];
            gA[i] = a.GT[base + i];
            gF[i] = a.GT[a.npos + base + i];
            gB[i] = a.GT[2 * a.npos + base + i];
            cA[i] = a.pta[base + i];
            cF[i] = a.ptcf[base + i];
            cB[i] = a.ptcb[base + i];
            sPar[i] = a.ppar[base + i];
            sObs[i] = a.obs_idx[base + i];
        }
        for (int i = tid; i < pm.nfl + pm.nbl + 2; i += T) sLvl[i] = lvl[i];
        for (int i = tid; i < 2 * pm.nfe; i += T) sFE[i] = FE[i];
        for (int i = tid; i < pm.nbe; i += T) sBK[i] = BKP[i];
        for (int i = tid; i < pm.nke; i += T) sKI[i] = KI[i];
    }
    __syncthreads();
    const int* blv = sLvl + pm.nfl + 1;
    int widest = 0;
    for (int l = 0; l < pm.nfl; l++) widest = max(widest, sLvl[l + 1] - sLvl[l]);
    for (int l = 0; l < pm.nbl; l++) widest = max(widest, blv[l + 1] - blv[l]);
    const bool one = __builtin_amdgcn_readfirstlane(widest) <= W;
    const bool sweeps = !one || wave == 0;      // this wave runs the sweeps
    const int s0 = one ? lane : tid, ss = one ? W : T;
    auto level_sync = [&]() {
        if (one)
            wave_lds_sync();
        else
            __syncthreads();
    };
    const double xs = d.crank ? 2.0 : 1.0;
    for (int j = a.n - 1; j >= 0; j--) {
        const long long t = (long long)a.t0 + j;
        const double* Vt = a.segV + (long long)j * a.npos;
        const double* St = a.segS + (long long)j * a.nS;
        for (int i = tid; i < NP; i += T) {
            const int o = sObs[i];
            double v = sV[i];
            if (o >= 0) v += a.dtrace[t * a.nobs + o];
            sV[i] = v;
            sY[i] = xs * v;
        }
        for (int blk = wave; blk < pm.nblk; blk += nw) {
            const BlockMeta bm = d.blocks[pm.block0 + blk];
            const int i = blk * W + lane;
            const CompOut o = adj_forward_local(d, d.prog + bm.prog, St + bm.dslot, lane, Vt[base + i]);
            sB[i] = o.rhs;
            sD[i] = o.diag;
        }
        __syncthreads();
        if (sweeps) {
            // forward elimination, as the step ran it
            for (int l = 0; l < pm.nfl; l++) {
                for (int e = sLvl[l] + s0; e < sLvl[l + 1]; e += ss) {
                    const int f0 = sFE[2 * e], kb = sFE[2 * e + 1];
                    const int p = f0 & 0xffff, nk = f0 >> 16;
                    double dp = sD[p], bp = sB[p];
                    for (int q = kb; q < kb + nk; q++) {
                        const int c = sKI[q];
                        const double tt = cA[c] / sD[c];
                        dp -= cF[c] * tt;
                        bp -= sB[c] * tt;
                    }
                    sD[p] = dp;
                    sB[p] = bp;
                }
                level_sync();
            }
            // back-substitution
            for (int l = 0; l < pm.nbl; l++) {
                for (int e = blv[l] + s0; e < blv[l + 1]; e += ss) {
                    const int kp = sBK[e];
                    const int c = kp & 0xffff, p = (kp >> 16) - 1;
                    double bv = sB[c];
                    if (p >= 0) bv -= cB[c] * sX[p];
                    sX[c] = bv / sD[c];
                }
                level_sync();
            }
            // back-substitution in reverse: children before parents
            for (int l = 0; l < pm.nfl; l++) {
                for (int e = sLvl[l] + s0; e < sLvl[l + 1]; e += ss) {
                    const int f0 = sFE[2 * e], kb = sFE[2 * e + 1];
                    const int p = f0 & 0xffff, nk = f0 >> 16;
                    double y = sY[p];
                    for (int q = kb; q < kb + nk; q++) {
                        const int c = sKI[q];
                        y -= sY[c] * cB[c] / sD[c];
                    }
                    sY[p] = y;
                }
                level_sync();
            }
            for (int i = s0; i < NP; i += ss) {
                const int p = sPar[i];
                if (p < -1) continue;
                const double dk = sD[i], xb = sY[i];
                const double bb = 0.0 + xb / dk;
                const double db = 0.0 - xb * sX[i] / dk;
                if (p >= 0) gB[i] = gB[i] - xb * sX[p] / dk;
                sY[i] = bb;
                sZ[i] = db;
            }
            level_sync();
            // forward elimination in reverse: parents before children
            for (int l = 0; l < pm.nbl; l++) {
                for (int e = blv[l] + s0; e < blv[l + 1]; e += ss) {
                    const int kp = sBK[e];
                    const int c = kp & 0xffff, p = (kp >> 16) - 1;
                    if (p < 0) continue;
                    const double dk = sD[c];
                    const double tt = cA[c] / dk;
                    const double bbp = sY[p], dbp = sZ[p];
                    sY[c] -= bbp * tt;
                    const double tbar = -bbp * sB[c] - dbp * cF[c];
                    gF[c] -= dbp * tt;
                    gA[c] += tbar / dk;
                    sZ[c] -= tbar * tt / dk;
                }
                level_sync();
            }
        }
        if (one) __syncthreads();
        // the programs in reverse
        for (int blk = wave; blk < pm.nblk; blk += nw) {
            const BlockMeta bm = d.blocks[pm.block0 + blk];
            const int i = blk * W + lane;
            if (lane < bm.lanes) {
                double vbar = adj_reverse_local(d, d.prog + bm.prog, St + bm.dslot, a.LS + bm.dslot, lane,
                                                Vt[base + i], sY[i], sZ[i]);
                // Crank-Nicolson: V^{t+1} = 2x - V^t
                if (d.crank) vbar -= sV[i];
                sV[i] = vbar;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < NP; i += T) {
        a.LV[base + i] = sV[i];
        a.GT[base + i] = gA[i];
        a.GT[a.npos + base + i] = gF[i];
        a.GT[2 * a.npos + base + i] = gB[i];
    }
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

hipError_t launch_adj_forest(const AdjPack& a, int max_pack_pos, hipStream_t s)
{
    if (a.d.npacks <= 0 || a.n <= 0) return hipSuccess;
    const unsigned lds = (unsigned)((size_t)ADJ_PACK_LDS_PER_POS * max_pack_pos + 16);
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
