// hsh_adjoint_device.h -- device code of the adjoint shared by the library's
// kernels (hsh_adjoint.hip) and the kernels generated at run time
// (hsh_jit.cpp): the reverse of a compartment's channel program and the pack
// layout's reverse step.  No includes: the run-time compiler takes it after
// hsh_devtypes.h and hsh_device.h.
#pragma once

namespace hsh {

constexpr int MAXG = 3;     // gates per channel the reverse program keeps (more are refused)

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
__device__ __forceinline__ CompOut adj_forward_local(const DevModel& d, const int* __restrict__ code, const double* __restrict__ Sb,
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
__device__ __forceinline__ double adj_reverse_local(const DevModel& d, const int* __restrict__ code, const double* __restrict__ Sb,
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

// The interpreter as the pack kernel's program reverse.
struct AdjInterp {
    static __device__ __forceinline__ double rev(const DevModel& d, const BlockMeta& bm, const double* Sb,
                                                 double* LSb, int lane, double Vm, double bb, double db)
    {
        return adj_reverse_local(d, d.prog + bm.prog, Sb, LSb, lane, Vm, bb, db);
    }
};

// Every reverse step of one checkpoint segment on the pack layout, one
// workgroup per pack (neurons are independent blocks of the system, so a
// pack's adjoint never reads another pack).  The pack's working arrays, its
// level schedule and its tree coefficients are in LDS for the whole launch
// (ADJ_PACK_LDS_PER_POS bytes per position).  Per step t, last first:
//   * dL/dV of the observed rows enters the adjoint of V^{t+1};
//   * the eliminated rhs and diag and the solution of step t, as the
//     recording launch computed them (prefetched one step ahead);
//   * the back-substitution in reverse, by height: a parent gathers its
//     kids' terms xb*tcb/dk in ascending row order -- the order in which
//     k_adj_elim's ascending rows subtract them -- and then every row's
//     adjoints of rhs and diag and the tcb gradient follow from its final
//     value;
//   * the forward elimination in reverse, by depth: a row reads its parent's
//     final adjoints of rhs and diag;
//   * each compartment's program in reverse (REV::rev: the interpreter
//     adj_reverse_local, or a straight-line function generated per program).
// Every value is computed by the operations of k_adj_elim in the same
// order (the recorded elimination is the one k_adj_elim recomputes), so the
// gradients equal the tile layout's bit for bit.  A pack whose
// levels are all at most one wave wide runs the sweeps in wave 0 alone (one
// wave's LDS accesses complete in order: no workgroup barrier per level);
// otherwise every wave takes part, with a barrier after each level.
constexpr int ADJ_PF = 4;

template <class REV>
__device__ __forceinline__ void body_adj_forest(const AdjPack& a, double* lds)
{
    const DevModel& d = a.d;
    const PackMeta pm = d.packs[blockIdx.x];
    const int NP = pm.nblk * W;
    const int T = blockDim.x, tid = threadIdx.x;
    const int nw = T >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & (W - 1);
    const long long base = (long long)pm.block0 * W;
    double* sV = lds;            // adjoint of V
    double* sB = sV + NP;        // eliminated rhs
    double* sD = sB + NP;        // eliminated diag
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
    // each step's recorded elimination, loaded into registers one step ahead
    // (at most ADJ_PF positions per thread; larger packs load in the step)
    const bool pf = NP <= ADJ_PF * T;
    double pb[ADJ_PF], pd[ADJ_PF], px[ADJ_PF];
    auto prefetch = [&](int j) {
        const long long o = (long long)j * a.npos + base;
#pragma unroll
        for (int q = 0; q < ADJ_PF; q++) {
            const int i = tid + q * T;
            if (i < NP) {
                pb[q] = a.segB[o + i];
                pd[q] = a.segD[o + i];
                px[q] = a.segX[o + i];
            }
        }
    };
    if (pf && a.n > 0) prefetch(a.n - 1);
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
        // the elimination of step t as the recording launch ran it:
        // eliminated rhs and diag, solution
        if (pf) {
#pragma unroll
            for (int q = 0; q < ADJ_PF; q++) {
                const int i = tid + q * T;
                if (i < NP) {
                    sB[i] = pb[q];
                    sD[i] = pd[q];
                    sX[i] = px[q];
                }
            }
            if (j > 0) prefetch(j - 1);
        } else {
            const long long o = (long long)j * a.npos + base;
            for (int i = tid; i < NP; i += T) {
                sB[i] = a.segB[o + i];
                sD[i] = a.segD[o + i];
                sX[i] = a.segX[o + i];
            }
        }
        __syncthreads();
        if (sweeps) {
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
                double vbar = REV::rev(d, bm, St + bm.dslot, a.LS + bm.dslot, lane, Vt[base + i], sY[i], sZ[i]);
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

}  // namespace hsh
