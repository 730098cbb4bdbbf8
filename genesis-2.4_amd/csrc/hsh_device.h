// hsh_device.h -- device code shared by the built-in kernels (hsh_kernels.hip)
// and the kernels libhsolve_hip specialises and compiles at run time
// (hsh_jit.cpp).  Include after hip_runtime.h and hsh_devtypes.h.
//
// Arithmetic follows the reference CPU code expression by expression
// (genesis/src/hines/hines_chip.c:157-600, hines_solve.c:140-296); compiled
// with -ffp-contract=off, every result rounds as the reference's does, and
// fp64 division is IEEE correctly rounded on gfx950.
#pragma once

namespace hsh {

struct CompOut {
    double rhs, diag;
};

// Table row and fraction for voltage Vm (hines_chip.c:174-203).  The
// reference keeps the row of the previous NEWVOLT compartment (xlo/xhi cache)
// and reuses it when Vm falls inside that row's [xvals[f], xvals[f+1]); the
// previous compartment's own row is recomputed here from its voltage.
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

// Gate state update from the interpolated B'/A' columns (hines_chip.c:425-434,
// :484-489): integrating gate X' = (X(2-B')+A')/B', instantaneous X = A'/B'.
__device__ __forceinline__ double gate_update(const double* __restrict__ vtab, int col, int ncols,
                                              double vipol, double s, bool instant)
{
    const double* tab = vtab + col;
    double X = tab[0];
    X = X + vipol * (tab[ncols] - X);
    double Y = tab[1];
    Y = Y + vipol * (tab[ncols + 1] - Y);
    if (!instant) return (s * (2.0 - X) + Y) / X;
    return Y / X;
}

// Gk *= X^k with the reference's multiplication sequences (hines_chip.c:490-509).
__device__ __forceinline__ double apply_power(double Gk, double X, int k)
{
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
    return Gk;
}

// SPIKE_OP (hines_chip.c:266-276): counter-1 every step; when Vm > thresh and
// the counter is spent, record an event and reload the counter.  With the
// synchans' lists on, the spike pass of the step has already run every
// SPIKE_OP (it reads only Vm, thresh and the counter), so here it does
// nothing.
__device__ __forceinline__ void spike_op(const DevModel& d, int* __restrict__ Ib, int is, int lane,
                                         double Vm, double thresh, bool active, int step)
{
    if (d.ev) return;
    int cnt = Ib[is * W + lane] - 1;
    if (Vm > thresh) {
        if (cnt <= 0) {
            if (active) {
                int q = atomicAdd(d.nspk, 1);
                if (q < d.spk_cap) {
                    d.spk[2 * q] = Ib[(is + 2) * W + lane];
                    d.spk[2 * q + 1] = step;
                }
            }
            cnt = Ib[(is + 1) * W + lane];
        }
    }
    Ib[is * W + lane] = cnt;
}

// EVENT of an hsolved synchan at simulation time `now`
// (newconn/synchan.c:232-300): a list that is not empty first re-times its
// head by the whole steps (int)((now - tle)/dt) since time_last_event; the
// event goes before the first event whose time field is >= its own (time
// fields relative to the event before); a new head sets the SYN2_OP counter
// (HSEVENT, hines.c:352-357); time_last_event = now.  Returns false when the
// list is full.
__device__ __forceinline__ bool ev_insert(int* __restrict__ lt, double* __restrict__ lw, int& n, int cap,
                                          int& cnt, double& tle, int T, double wdt, double now, double dt)
{
    if (n >= cap) return false;
    if (n == 0) {
        lt[0] = T;
        lw[0] = wdt;
        n = 1;
        cnt = T;
    } else {
        lt[0] -= (int)((now - tle) / dt);
        int i = 0;
        for (; i < n; i++) {
            if (T <= lt[i]) break;
            T -= lt[i];
        }
        if (i < n) {
            lt[i] -= T;
            for (int j = n; j > i; j--) {
                lt[j] = lt[j - 1];
                lw[j] = lw[j - 1];
            }
        }
        lt[i] = T;
        lw[i] = wdt;
        n++;
        if (i == 0) cnt = T;
    }
    tle = now;
    return true;
}

// SYN2_OP (hines_chip.c:511-550): X *= xconst2; if the counter is 0 the
// synchan fires (h_dosynchan, hines_conn.c:60-89; HPROCESS,
// newconn/synchan.c:305-358): the events at the front whose time field is 0
// (the head's forced to 0) are popped, weight/dt summed into the float
// activation, X += dt*activation*stablist[k+3], and the counter takes the new
// head's time or -1; the counter decrements; Y = X*yconst1 + Y*yconst2.
// With the lists on, the EVENTs this step's spikes send to the synchan run
// around it in ops[] order: those of SPIKE_OPs before the SYN2_OP first (they
// may re-arm the counter), the rest after the decrement.  Returns Y (the
// caller multiplies it into Gk and adds the current).  Padding lanes repeat
// lane 0's synchan and must not touch its list.
__device__ __forceinline__ double syn2_op(const DevModel& d, double* __restrict__ Sb, int* __restrict__ Ib, int c,
                                          int k, int is, int lane, double X0, double Y0, bool active, int step)
{
    double X = X0 * d.stab[k];
    if (d.ev && active) {
        const int sid = Ib[(is + 1) * W + lane];
        int cnt = Ib[is * W + lane];
        const int par = step & 1;
        int* __restrict__ narr = d.narr + par * d.ev_nsyn;
        const int na = narr[sid];
        if (na > 0 || cnt == 0) {
            const int j0 = d.in_off[sid], mid = d.in_mid[sid];
            const int off = d.lst_off[sid], cap = d.lst_cap[sid];
            int* __restrict__ lt = d.lst_t + off;
            double* __restrict__ lw = d.lst_w + off;
            int* __restrict__ ar = d.arr + par * d.ev_nconn + j0;
            int n = d.lst_n[sid];
            double tle = d.tle[sid];
            const double now = d.tnow[par];
            bool full = false;
            // the arrivals in ops[] order of their sources (slot order)
            for (int a = 1; a < na; a++) {
                const int v = ar[a];
                int b = a - 1;
                while (b >= 0 && ar[b] > v) {
                    ar[b + 1] = ar[b];
                    b--;
                }
                ar[b + 1] = v;
            }
            int a = 0;
            for (; a < na && ar[a] < mid; a++)
                full |= !ev_insert(lt, lw, n, cap, cnt, tle, d.in_T[ar[a]], d.in_wdt[ar[a]], now, d.dt);
            if (cnt == 0) {
                float act = 0.0f;
                if (n > 0) {
                    lt[0] = 0;
                    int i = 0;
                    while (i < n && lt[i] == 0) {
                        act = (float)((double)act + lw[i]);
                        i++;
                    }
                    for (int q = i; q < n; q++) {
                        lt[q - i] = lt[q];
                        lw[q - i] = lw[q];
                    }
                    n -= i;
                    if (n > 0) tle = now;
                }
                X += d.dt * (double)act * d.stab[k + 3];
                cnt = n > 0 ? lt[0] : -1;
            }
            cnt -= 1;
            for (; a < na; a++)
                full |= !ev_insert(lt, lw, n, cap, cnt, tle, d.in_T[ar[a]], d.in_wdt[ar[a]], now, d.dt);
            if (na > 0) narr[sid] = 0;
            d.lst_n[sid] = n;
            d.tle[sid] = tle;
            if (full) atomicOr(d.everr, 2);
        } else {
            cnt -= 1;
        }
        Ib[is * W + lane] = cnt;
    } else {
        Ib[is * W + lane] -= 1;
    }
    Sb[c * W + lane] = X;
    const double Y = X * d.stab[k + 1] + Y0 * d.stab[k + 2];
    Sb[(c + 1) * W + lane] = Y;
    return Y;
}

// Concentration table row and fraction (hines_chip.c:236-252), looked up
// afresh (the reference's ylo/yhi row cache gives the same row except within
// rounding of a breakpoint).
__device__ __forceinline__ void newconc(const DevModel& d, double concen, int& cfilo, double& cipol)
{
    double v = (concen - d.ymin) * d.invdy;
    if (v < 0.0) v = 0.0;
    else if (v > d.ydivs) v = d.ydivs;
    int f = (int)v;
    f = f < 0 ? 0 : (f > d.ydivs ? d.ydivs : f);
    cfilo = f;
    cipol = v - (double)f;
}

// Table lookup of the IPOL ops (hines_chip.c:419-458, :567-584).  which:
// 0 = 1-D on the voltage row, 1 = 1-D on the concentration row, 2 = 2-D over
// voltage and concentration, 3 = 2-D over two concentrations from the table
// register (the position the previous lookup left, hines_chip.c:569).  The
// table register moves to one past the column read, as the reference's does.
__device__ __forceinline__ void ipol_xy(const DevModel& d, int which, int col, const double* vtab,
                                        const double* ctab, const double*& tabr, int cfilo, double vipol,
                                        double cipol, double ipol, double& X, double& Y)
{
    const int nc = d.ncols;
    const double* tb;
    if (which <= 1) {
        tb = (which == 0 ? vtab : ctab) + col;
        const double f = which == 0 ? vipol : cipol;
        X = tb[0];
        X = X + f * (tb[nc] - X);
        Y = tb[1];
        Y = Y + f * (tb[nc + 1] - Y);
    } else {
        long long off = (long long)((which == 2 ? vtab : tabr) - d.tab) + 2LL * cfilo + col;
        if (which == 3) off = off < 0 ? 0 : (off > d.tab_max ? d.tab_max : off);
        tb = d.tab + off;
        const double f = which == 2 ? vipol : ipol;
        const double* tn = tb + nc;
        X = (1.0 - f) * ((1.0 - cipol) * tb[0] + cipol * tb[2]) + f * ((1.0 - cipol) * tn[0] + cipol * tn[2]);
        Y = (1.0 - f) * ((1.0 - cipol) * tb[1] + cipol * tb[3]) + f * ((1.0 - cipol) * tn[1] + cipol * tn[3]);
    }
    tabr = tb + 1;
}

// GHK_OP (hines_chip.c:334-365): slope conductance and the Ek that gives the
// GHK current at Vm.  exp() is the device library's.
struct GhkOut {
    double Ek, slope;
};
__device__ __forceinline__ GhkOut ghk_eval(double Vm, double Cin, double Cout, double zFRT, double zF)
{
    double GHK = zFRT;
    const double X = Vm * GHK;
    const double Y = exp(-X);
    const double E = zF;
    double Ik = E * (Cin - (Y * Cout));
    double G;
    if (fabs(X) < 1.0e-6) {
        G = 1.0 - 0.5 * X;
        GHK = 0.5 * GHK * E * (Cin + Cout);
        Ik = -Ik / G;
    } else {
        G = 1.0 - Y;
        GHK = GHK * (Ik / G - E * GHK * (Cin - Cout) * Vm * Y / (G * G));
        Ik = -X * Ik / G;
    }
    return GhkOut{(Ik + GHK * Vm) / GHK, GHK};
}
// The interpreter calls it out of line (its exp would raise the register
// count of every interpreted program); results come back in registers.
__device__ __noinline__ GhkOut ghk_call(double Vm, double Cin, double Cout, double zFRT, double zF)
{
    return ghk_eval(Vm, Cin, Cout, zFRT, zF);
}

// Pool flux registers with a run-time pool number (interpreter): selects over
// constant indices keep them in VGPRs, where an indexed array or a reference
// into one would live in scratch memory.
struct Flux {
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0;
    __device__ __forceinline__ double get(int p) const
    {
        return p == 0 ? f0 : (p == 1 ? f1 : (p == 2 ? f2 : f3));
    }
    __device__ __forceinline__ void set(int p, double v)
    {
        if (p == 0) f0 = v;
        else if (p == 1) f1 = v;
        else if (p == 2) f2 = v;
        else f3 = v;
    }
};

// givals of one current (hines_chip.c:536-564): Gk, Ik, Ek at gi, gi+1,
// gi+2 (chanmode 5: Gk and Ik per gmax).
__device__ __forceinline__ void giv_put3(const DevModel& d, int& gi, double Gk, double Ik, double Ek, double gmax)
{
    if (d.chanmode5) {
        d.giv[gi] = Gk / gmax;
        d.giv[gi + 1] = Ik / gmax;
    } else {
        d.giv[gi] = Gk;
        d.giv[gi + 1] = Ik;
    }
    d.giv[gi + 2] = Ek;
    gi += 3;
}

// Mg block of a synaptic conductance (hines_chip.c:531-547): X = KMg_A
// exp(Vm / KMg_B), Gk *= X / (X + [Mg]).  exp() is the device library's.
__device__ __forceinline__ double mgblock(double Gk, double Vm, double A, double invB, double Mg)
{
    const double X = A * exp(Vm * invB);
    return Gk * (X / (X + Mg));
}

// Ca_concen pool (hines_conc_chip.c:196-207): C = base + ((C - base)(2 - i')
// + influx * B') / i'.
__device__ __forceinline__ double caconc(double C, double base, double itau, double influx, double Bp)
{
    return base + ((C - base) * (2.0 - itau) + (influx * Bp)) / itau;
}

// ---------------------------------------------------------------------------
// A compartment's concentration program and diffusion systems
// (hines_conc_chip.c:59-222, hines_conc.c:39-396), run at the end of the
// program of the compartment that owns them.  They read and write the
// compartment's slots in memory (a shell's concentration, diagonal and result
// change within the step), with the same expressions in the same order as
// the reference.

// C^(hill-1) as the Hill pumps form it (hines_conc_chip.c:126-135)
__device__ __forceinline__ double hill_pow(int hill, double X)
{
    if (hill == 2) return X;
    if (hill == 3) return X * X;
    if (hill == 4) return X * X * X;
    return 1.0;
}

// One diffusion system.  Slots: concentrations cs0 + r (row r = shell*dw +
// row in shell, shell 0 innermost), covals0 values cv0 + l in the order they
// are used, buffer rates bv0 + 3*buffer.  Per-step arrays, TS apart (W: slots
// in memory, the interpreter; 1: a thread's own arrays, generated code): the
// shells' codiag and coresult, the eliminated upper entries (dw per row) and
// right-hand sides.  Shells alone (dw 1): tridiagonal, Crank-Nicolson as
// C <- 2x - C; with buffers: the banded Crank-Nicolson system, whose
// diffusible buffer rows keep the reference's second value of their upper
// entry (hines_conc.c:250).  The diagonals used are set back to 1.
template <int TS>
__device__ __forceinline__ void diff_solve_t(double* __restrict__ Sb, int lane, int ms, int dw, int nd, int cs0, int cv0,
                                    int bv0, double* __restrict__ dgp, double* __restrict__ crp,
                                    double* __restrict__ xp, double* __restrict__ rp, int crank)
{
#define DS(c) Sb[(c) * W + lane]
#define CN(r) DS(cs0 + (r))
#define CV(i) xp[(i) * TS]
#define RS(i) rp[(i) * TS]
#define K0(i) DS(cv0 + (i))
#define BV(i) DS(bv0 + (i))
#define DG(m) dgp[(m) * TS]
#define CR(m) crp[(m) * TS]
    int l = 0;
    if (dw == 1) {
        double up = K0(l++);
        double dg = up + DG(0);
        DG(0) = 1.0;
        RS(0) = (CN(0) + CR(0)) / dg;
        int j;
        for (j = 1; j < ms - 1; j++) {
            const double e = -up / dg;
            CV(j) = e;
            const double dn = K0(l);
            up = K0(l + 1);
            l += 2;
            dg = DG(j) + dn * (1.0 + e) + up;
            DG(j) = 1.0;
            RS(j) = (CN(j) + CR(j) + dn * RS(j - 1)) / dg;
        }
        const double e = -up / dg;
        CV(j) = e;
        const double dn = K0(l++);
        dg = DG(j) + dn * (1.0 + e);
        DG(j) = 1.0;
        double x = (CN(j) + CR(j) + dn * RS(j - 1)) / dg;
        if (crank) {
            CN(j) = x + x - CN(j);
            for (int r = j - 1; r >= 0; r--) {
                x = RS(r) - CV(r + 1) * x;
                CN(r) = x + x - CN(r);
            }
        } else {
            CN(j) = x;
            for (int r = j - 1; r >= 0; r--) {
                x = RS(r) - CV(r + 1) * x;
                CN(r) = x;
            }
        }
        return;
    }
    const int sw = dw * dw;
    double co = CN(dw - 1);
    for (int m = 0; m < ms; m++) {
        const int rS = m * dw + dw - 1, xS = rS * dw, xP = xS - sw, rP = rS - dw;
        const bool last = (m == ms - 1);
        double dgc = DG(m);
        DG(m) = 1.0;
        double rsc = CR(m);
        for (int n = 0; n < nd; n++) {
            const int r = m * dw + n, nb = 3 * n;
            double dg = BV(nb) * co + BV(nb + 1);
            CV(xS + n) = dg;
            const double t = BV(nb) * CN(r);
            dgc += t;
            dg = 1.0 / (1.0 + dg);
            CV(r * dw) = t * dg;
            const double t2 = -BV(nb + 1) * CN(r) + K0(l++);
            rsc += t2;
            RS(r) = (CN(r) + t2) * dg;
        }
        for (int n = nd; n < dw - 1; n++) {
            const int r = m * dw + n, nb = 3 * n;
            double dg = BV(nb) * co + BV(nb + 1);
            CV(xS + n) = dg;
            const double t = BV(nb) * CN(r);
            dgc += t;
            double t2 = -BV(nb + 1) * CN(r) + BV(nb + 2);
            rsc += t2;
            if (m) {
                const double nw = K0(l) * CV((r - dw) * dw);
                if (!last) {
                    dg += 1.0 + K0(l + 1) + K0(l) * (1.0 + CV((r - dw) * dw + 1)) - nw * CV(xP + n);
                    CV(r * dw + 1) = -K0(l + 1) / dg;
                    t2 += K0(l) * CN(r - dw) + (1.0 - K0(l) - K0(l + 1)) * CN(r) + K0(l + 1) * CN(r + dw);
                } else {
                    dg += 1.0 + K0(l) * (1.0 + CV((r - dw) * dw + 1)) - nw * CV(xP + n);
                    t2 += K0(l) * CN(r - dw) + (1.0 - K0(l)) * CN(r);
                }
                RS(r) = (t2 + K0(l) * RS(r - dw) - nw * RS(rP)) / dg;
                l += last ? 1 : 2;
            } else {
                dg += 1.0 + K0(l);
                CV(r * dw + 1) = -K0(l) / dg;
                RS(r) = (t2 + (1.0 - K0(l)) * CN(r) + K0(l) * CN(r + dw)) / dg;
                l++;
            }
            CV(r * dw) = t / dg;
        }
        if (m) {
            if (!last) {
                dgc += K0(l + 1) + K0(l) * (1.0 + CV(xP + dw - 1));
                rsc += K0(l) * (CN(rS - dw) + RS(rP)) + (1.0 - K0(l) - K0(l + 1)) * co;
                co = CN(rS + dw);
                rsc += K0(l + 1) * co;
            } else {
                dgc += K0(l) * (1.0 + CV(xP + dw - 1));
                rsc += K0(l) * (CN(rS - dw) + RS(rP)) + (1.0 - K0(l)) * co;
            }
        } else {
            dgc += K0(l);
            rsc += (1.0 - K0(l)) * co;
            co = CN(rS + dw);
            rsc += K0(l) * co;
        }
        for (int n = 0; n < nd; n++) {
            const int rb = m * dw + n;
            dgc -= CV(xS + n) * CV(rb * dw);
            rsc -= CV(xS + n) * RS(rb);
        }
        for (int n = nd; n < dw - 1; n++) {
            const int rb = m * dw + n;
            const double nw = m ? CV(xS + n) + K0(l) * CV(xP + n) : CV(xS + n);
            dgc -= nw * CV(rb * dw);
            rsc -= nw * RS(rb);
            if (!last) CV(xS + n) *= -CV(rb * dw + 1);
        }
        if (!last) {
            for (int n = nd; n < dw - 1; n++) CV(xS + n) /= dgc;
            if (m) {
                CV(xS + dw - 1) = -K0(l + 1) / dgc;
                l += 2;
            } else {
                CV(xS + dw - 1) = -K0(l) / dgc;
                l++;
            }
            RS(rS) = rsc / dgc;
        } else {
            co = rsc / dgc;
            CN(rS) = co;
            l++;
        }
    }
    for (int m = ms - 1; m >= 0; m--) {
        const int rS = m * dw + dw - 1, xS = rS * dw;
        if (m < ms - 1) {
            co = RS(rS) - CV(xS + dw - 1) * CN(rS + dw);
            for (int n = nd; n < dw - 1; n++) co -= CV(xS + n) * CN(rS + n + 1);
            CN(rS) = co;
            for (int n = dw - 2; n >= nd; n--) {
                const int rb = m * dw + n;
                CN(rb) = RS(rb) - CV(rb * dw) * co - CV(rb * dw + 1) * CN(rb + dw);
            }
        } else {
            for (int n = dw - 2; n >= nd; n--) {
                const int rb = m * dw + n;
                CN(rb) = RS(rb) - CV(rb * dw) * co;
            }
        }
        for (int n = nd - 1; n >= 0; n--) {
            const int rb = m * dw + n;
            CN(rb) = RS(rb) - CV(rb * dw) * co;
        }
    }
#undef DS
#undef CN
#undef CV
#undef RS
#undef K0
#undef BV
#undef DG
#undef CR
}

// The concentration ops of hines_conc_chip.c (a: first value slot; dg: the
// shell's codiag, wherever it is kept)
__device__ __forceinline__ void conc_fluxconc(const double* __restrict__ Sb, int lane, double& csum, int a,
                                              double flux)
{
    csum += Sb[a * W + lane] + flux * Sb[(a + 1) * W + lane];
}
__device__ __forceinline__ void conc_ctaupump(const double* __restrict__ Sb, int lane, double& csum, int a, double& dg)
{
    csum += Sb[a * W + lane];
    dg += Sb[(a + 1) * W + lane];
}
__device__ __forceinline__ void conc_taupump(const double* __restrict__ Sb, int lane, double& csum, int a, double& dg,
                                             double dt, double Vm)
{
    const double itau = dt / (Sb[a * W + lane] + Sb[(a + 1) * W + lane] *
                                                      exp((Vm - Sb[(a + 2) * W + lane]) * Sb[(a + 3) * W + lane]));
    csum += Sb[(a + 4) * W + lane] * itau;
    dg += itau;
}
__device__ __forceinline__ void conc_mmpump(const double* __restrict__ Sb, int lane, int a, double C, double& dg)
{
    dg += Sb[a * W + lane] / (C + Sb[(a + 1) * W + lane]);
}
__device__ __forceinline__ void conc_hillpump(const double* __restrict__ Sb, int lane, int a, double X, double& dg,
                                              int hill)
{
    double c = hill_pow(hill, X);
    c = c / (c * X + Sb[a * W + lane]);
    dg += Sb[(a + 1) * W + lane] * c;
}
__device__ __forceinline__ void conc_phillpump(const double* __restrict__ Sb, int lane, double& csum, int a, double X,
                                               double& dg, int hill)
{
    double c = hill_pow(hill, X);
    c = c / (c * X + Sb[a * W + lane]);
    dg += Sb[(a + 1) * W + lane] * c;
    csum += Sb[(a + 2) * W + lane] * c * X;
}
__device__ __forceinline__ void conc_buffer(double* __restrict__ Sb, int lane, double& csum, int a, int cb, int cp)
{
    const double X = Sb[a * W + lane];
    const double B = Sb[cb * W + lane];
    const double c = ((1.0 - X) * B + Sb[(a + 1) * W + lane]) / (1.0 + X + Sb[(a + 2) * W + lane] * Sb[cp * W + lane]);
    csum += c - B;
    Sb[cb * W + lane] = c;
}

// ---------------------------------------------------------------------------
// Generic micro-op interpreter.  With REG the compartment's whole slot vector
// is fetched with back-to-back coalesced loads before the program runs (one
// memory latency per compartment instead of one per micro-op) and indexed
// with the wave-uniform slot cursor, which gfx950 does in VGPRs
// (s_set_gpr_idx).  A slot is never read after the program writes it (the
// cursor only moves forward), so writes go to memory only and the register
// copy stays read-only.  Without REG each slot is read when its micro-op runs.
// Without EXT only the voltage-gated and synaptic ops are compiled (END,
// NEWVOLT, CHAN, CHAN_EK, IPOL1V, ADDCURR, SPIKE, SYN2, GIV0): the
// concentration, 2-D table, GHK and Mg-block cases would otherwise set the
// register budget, and with it the occupancy, of every program; the host
// picks EXT when any program holds another op.
constexpr int MAXS = 12;
// rows of tree coefficients body_batched's one-wave elimination loads ahead
constexpr int BCH = 16;

template <bool REG, bool EXT>
__device__ __forceinline__ CompOut run_prog(const DevModel& d, const int* __restrict__ code,
                                            double* __restrict__ Sb, int* __restrict__ Ib, int lane,
                                            double Vm, bool pred_ok, double Vp, bool active, int step)
{
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
#define SGET(c) (REG ? s[(c)] : Sb[(c) * W + lane])
#define SPUT(c, v) (Sb[(c) * W + lane] = (v))
    double sumgchan = 0.0;
    double ichan = SGET(0) + SGET(1);
    double Gk = 0.0, Ek = 0.0, Ik = 0.0;
    double vipol = 0.0, cipol = 0.0;
    Flux flux;
    double csum = 0.0;
    int cfilo = 0;
    double ipol = 0.0, ghk = 0.0;
    const double* tabr = d.tab;
    const double* vtab = d.tab;
    const double* ctab = d.tab;
    const int ncols = d.ncols;
    // givals: cursor, membrane current, gmax register (chanmode 5)
    const bool giv = d.giv != nullptr && active;
    int gi = 0;
    double Im = 0.0, gmax = 0.0;
    int c = 2;
    int pc = PROG_HEADER;
    for (;;) {
        const int op = code[pc++];
        if (op == U_IPOL1V) {
            const int col = code[pc];
            const int k = code[pc + 1];
            pc += 2;
            const double X = gate_update(vtab, col, ncols, vipol, SGET(c), k < 0);
            tabr = vtab + col + 1;
            SPUT(c, X);
            c++;
            Gk = apply_power(Gk, X, k < 0 ? -k : k);
        } else if (op == U_CHAN_EK) {
            Ek = SGET(c);
            gmax = Gk = SGET(c + 1);
            c += 2;
        } else if (op == U_ADDCURR) {
            sumgchan += Gk;
            ichan += Ek * Gk;
            Ik = Gk * (Ek - Vm);
            if (giv) {
                giv_put3(d, gi, Gk, Ik, Ek, gmax);
                Im += Ik;
            }
        } else if (op == U_NEWVOLT) {
            int filo;
            newvolt(d, Vm, pred_ok, Vp, filo, vipol);
            vtab = d.tab + (long long)filo * ncols;
        } else if (op == U_CHAN) {
            gmax = Gk = SGET(c);
            c += 1;
        } else if (op == U_SPIKE) {
            const int is = code[pc++];
            spike_op(d, Ib, is, lane, Vm, SGET(c), active, step);
            c += 1;
        } else if (op == U_SYN2 || (EXT && op == U_SYN2MG)) {
            const int k = code[pc], is = code[pc + 1];
            pc += 2;
            Gk *= syn2_op(d, Sb, Ib, c, k, is, lane, SGET(c), SGET(c + 1), active, step);
            c += 2;
            if (EXT && op == U_SYN2MG) {
                if (giv) giv_put3(d, gi, Gk, Gk * (Ek - Vm), Ek, gmax);   // unblocked
                Gk = mgblock(Gk, Vm, SGET(c), SGET(c + 1), SGET(c + 2));
                c += 3;
            }
            sumgchan += Gk;
            ichan += Ek * Gk;
            Ik = Gk * (Ek - Vm);
            if (giv) {
                giv_put3(d, gi, Gk, Ik, Ek, gmax);
                Im += Ik;
            }
        } else if (EXT && op == U_CONCVAL) {
            flux.set(code[pc++], 0.0);
        } else if (EXT && (op == U_NEWCONC1 || op == U_DOMAIN)) {
            double concen;
            if (op == U_NEWCONC1) {
                concen = SGET(code[pc++]);
            } else {
                concen = SGET(c) * Ik;
                c++;
            }
            newconc(d, concen, cfilo, cipol);
            ctab = d.tab + (long long)cfilo * ncols;
        } else if (EXT && op == U_NEWCONC2) {
            tabr = ctab;
            ipol = cipol;
            newconc(d, SGET(code[pc++]), cfilo, cipol);
            ctab = d.tab + (long long)cfilo * ncols;
        } else if (EXT && op == U_IPOL1C) {
            const int col = code[pc];
            const int k = code[pc + 1];
            pc += 2;
            const double X = gate_update(ctab, col, ncols, cipol, SGET(c), k < 0);
            tabr = ctab + col + 1;
            SPUT(c, X);
            c++;
            Gk = apply_power(Gk, X, k < 0 ? -k : k);
        } else if (EXT && op == U_IPOLG) {
            const int which = code[pc], col = code[pc + 1], k = code[pc + 2];
            pc += 3;
            double X, Y;
            ipol_xy(d, which, col, vtab, ctab, tabr, cfilo, vipol, cipol, ipol, X, Y);
            X = (k > 0) ? (SGET(c) * (2.0 - X) + Y) / X : Y / X;
            SPUT(c, X);
            c++;
            Gk = apply_power(Gk, X, k < 0 ? -k : k);
        } else if (EXT && (op == U_TABCURR || op == U_TABGHK)) {
            const int col = code[pc];
            const int which = code[pc + 1];
            pc += 2;
            double X, Y;
            ipol_xy(d, which, col, vtab, ctab, tabr, cfilo, vipol, cipol, ipol, X, Y);
            Ek = Y;
            if (op == U_TABGHK) {
                ghk = X;
            } else {
                Gk *= X;
                sumgchan += Gk;
                ichan += Ek * Gk;
                Ik = Gk * (Ek - Vm);
                if (giv) {
                    giv_put3(d, gi, Gk, Ik, Ek, gmax);
                    Im += Ik;
                }
            }
        } else if (EXT && op == U_SAMEGHK) {
            gmax = SGET(c);
            Gk = gmax * ghk;
            c++;
        } else if (op == U_GIV0) {
            const int rs = code[pc], is = code[pc + 1];
            pc += 2;
            if (giv) {
                gi = Ib[is * W + lane];
                Im = SGET(0) - Vm / SGET(rs);
                d.giv[gi + 1] = Im;
                gi += 2;
            }
        } else if (EXT && op == U_FLUX) {
            const int p = code[pc++];
            flux.set(p, flux.get(p) + Ik);
        } else if (EXT && op == U_SCALEFLUX) {
            const int p = code[pc++];
            flux.set(p, flux.get(p) + SGET(c) * Ik);
            c++;
        } else if (EXT && op == U_NERNST) {
            const int ci = code[pc], co = code[pc + 1];
            pc += 2;
            Ek = log(SGET(co) / SGET(ci)) * SGET(c);
            gmax = Gk = SGET(c + 1);
            c += 2;
        } else if (EXT && op == U_GHK) {
            const int ci = code[pc], co = code[pc + 1];
            pc += 2;
            const GhkOut g = ghk_call(Vm, SGET(ci), SGET(co), SGET(c), SGET(c + 1));
            Ek = g.Ek;
            ghk = g.slope;
            gmax = SGET(c + 2);
            Gk = gmax * g.slope;
            c += 3;
        } else if (EXT && op == U_POOLCA) {
            const int p = code[pc], cs = code[pc + 1], ks = code[pc + 2];
            pc += 3;
            SPUT(cs, caconc(SGET(cs), SGET(ks), SGET(ks + 1), flux.get(p), SGET(ks + 2)));
        } else if (EXT && op == U_ELMMPUMP) {
            const int cs = code[pc], dg = code[pc + 1];
            pc += 2;
            const double C = Sb[cs * W + lane];
            const double X = 1.0 / (C + SGET(c));
            Ik = SGET(c + 1) * C * X;
            ichan += Ik;
            if (giv) {
                d.giv[gi] = d.chanmode5 ? Ik / SGET(c + 1) : Ik;
                gi += 1;
            }
            Im += Ik;
            if (dg >= 0) Sb[dg * W + lane] += SGET(c + 2) * X;
            c += 3;
        } else if (EXT && op == U_CZERO) {
            csum = 0.0;
        } else if (EXT && op == U_CSTORE) {
            Sb[code[pc++] * W + lane] = csum;
        } else if (EXT && op == U_FLUXCONC) {
            const int a = code[pc], p = code[pc + 1];
            pc += 2;
            conc_fluxconc(Sb, lane, csum, a, p >= 0 ? flux.get(p) : 0.0);
        } else if (EXT && op == U_CTAUPUMP) {
            conc_ctaupump(Sb, lane, csum, code[pc], Sb[code[pc + 1] * W + lane]);
            pc += 2;
        } else if (EXT && op == U_TAUPUMP) {
            conc_taupump(Sb, lane, csum, code[pc], Sb[code[pc + 1] * W + lane], d.dt, Vm);
            pc += 2;
        } else if (EXT && op == U_MMPUMP) {
            conc_mmpump(Sb, lane, code[pc], Sb[code[pc + 1] * W + lane], Sb[code[pc + 2] * W + lane]);
            pc += 3;
        } else if (EXT && op == U_HILLPUMP) {
            conc_hillpump(Sb, lane, code[pc], Sb[code[pc + 1] * W + lane], Sb[code[pc + 2] * W + lane], code[pc + 3]);
            pc += 4;
        } else if (EXT && op == U_PHILLPUMP) {
            conc_phillpump(Sb, lane, csum, code[pc], Sb[code[pc + 1] * W + lane], Sb[code[pc + 2] * W + lane],
                           code[pc + 3]);
            pc += 4;
        } else if (EXT && op == U_BUFFER) {
            conc_buffer(Sb, lane, csum, code[pc], code[pc + 1], code[pc + 2]);
            pc += 3;
        } else if (EXT && op == U_POOL) {
            Sb[code[pc] * W + lane] += csum;
            pc += 1;
        } else if (EXT && op == U_DIFF) {
            diff_solve_t<W>(Sb, lane, code[pc], code[pc + 1], code[pc + 2], code[pc + 3], code[pc + 4],
                            code[pc + 5], Sb + code[pc + 6] * W + lane, Sb + code[pc + 7] * W + lane,
                            Sb + code[pc + 8] * W + lane, Sb + code[pc + 9] * W + lane, d.crank);
            pc += 10;
        } else {  // U_END
            if (giv) d.giv[gi] = Im;
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

// Runner policy: the interpreter.  (Generated modules supply their own runner
// with one straight-line function per template.)
template <bool REG, bool EXT>
struct InterpRun {
    static __device__ __forceinline__ CompOut run(const DevModel& d, const BlockMeta& bm, int lane,
                                                  double Vm, bool pok, double Vp, bool active,
                                                  int step)
    {
        return run_prog<REG, EXT>(d, d.prog + bm.prog, d.S + bm.dslot, d.I + bm.islot, lane, Vm, pok, Vp,
                             active, step);
    }
};

// ---------------------------------------------------------------------------
// Kernel bodies (the __global__ wrappers pick the runner).

// Plain drop-in, part (a): one wave per block, four blocks per workgroup.
template <class RUN>
__device__ __forceinline__ void body_channel(const DevModel& d)
{
    int b = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    b = __builtin_amdgcn_readfirstlane(b);
    if (b >= d.nblocks) return;
    const int lane = threadIdx.x & (W - 1);
    const BlockMeta bm = d.blocks[b];
    const int pos = b * W + lane;
    const double Vm = d.V[pos];
    const int pp = d.pred[pos];
    const double Vp = pp >= 0 ? d.V[pp] : 0.0;
    const int step = d.step[bm.tile];
    CompOut o = RUN::run(d, bm, lane, Vm, pp >= 0, Vp, lane < bm.lanes, step);
    d.R[(long long)b * 2 * W + lane] = o.rhs;
    d.R[(long long)b * 2 * W + W + lane] = o.diag;
}

// Batched persistent kernel for small trees (rows <= 4*RPW): four waves per
// tile, wave w owns rows w*RPW .. w*RPW+RPW-1.  Each owner keeps its rows'
// tree coefficients in registers for the whole launch, and the elimination
// walks the rows in order -- the owner of row r folds it into its parent; the
// waves meet at a barrier where the next row belongs to another wave -- so
// the tile needs only 24 bytes of LDS per compartment and every tile of the
// workload is resident at once.
// CHAN = false is the plain drop-in's elimination of one step: rhs and diag
// come from results[] as the channel kernel left them, and the same owners
// and barriers run the sweep (four waves per tile instead of one).
template <int RPW, class RUN, bool CHAN = true>
__device__ __forceinline__ void body_batched_r(const DevModel& d, int K, double* lds)
{
    constexpr int NW = 4;
    constexpr int RMAX = NW * RPW;
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
    // the rows' block records are the same every step: read once (scalar
    // loads, the row is wave-uniform) instead of once per row per step
    BlockMeta bms[RPW];
#pragma unroll
    for (int i = 0; i < RPW; i++) {
        const int r = wave * RPW + i;
        if (r < R) bms[i] = d.blocks[b0 + r];
    }
    // every row's loads first, then the LDS stores of V (a store right
    // after its row's loads would wait for them, one latency per row)
    double v0[RPW], rb0[RPW], rd0[RPW];
    int pp[RPW];
#pragma unroll
    for (int i = 0; i < RPW; i++) {
        const int r = wave * RPW + i;
        par[i] = -1;
        pp[i] = -1;
        ta[i] = tcf[i] = tcb[i] = v0[i] = rb0[i] = rd0[i] = 0.0;
        if (r < R) {
            const long long q = (long long)(b0 + r) * W + lane;
            par[i] = d.par[q];
            ta[i] = d.ta[q];
            tcf[i] = d.tcf[q];
            tcb[i] = d.tcb[q];
            v0[i] = d.V[q];
            if (CHAN) {
                pp[i] = d.pred[q];
            } else {
                rb0[i] = d.R[(long long)(b0 + r) * 2 * W + lane];
                rd0[i] = d.R[(long long)(b0 + r) * 2 * W + W + lane];
            }
        }
    }
    double droot_in = 0.0;     // CHAN = false: the last row's diagonal as results[] holds it
#pragma unroll
    for (int i = 0; i < RPW; i++) {
        const int r = wave * RPW + i;
        pidx[i] = -1;
        if (r < R) {
            LV[r * W + lane] = v0[i];
            if (!CHAN) {
                LB[r * W + lane] = rb0[i];
                LD[r * W + lane] = rd0[i];
                if (r == R - 1) droot_in = rd0[i];
            }
            if (pp[i] >= 0 && (pp[i] >> 6) >= b0 && (pp[i] >> 6) < b0 + R)
                pidx[i] = ((pp[i] >> 6) - b0) * W + (pp[i] & (W - 1));
        }
    }
    const bool is_last = ((long long)(b0 + R - 1) * W + lane) == (long long)d.last_pos;
    const int step0 = d.step[t];
    __syncthreads();
    for (int k = 0; k < K; k++) {
        // the last row's diagonal as the channel program leaves it: its
        // owner keeps it (other waves may fold kids into it right after the
        // barrier) and puts it back after the sweep
        double droot0 = droot_in;
#pragma unroll 1
        for (int i = 0; i < RPW && CHAN; i++) {
            const int r = wave * RPW + i;
            if (r < R) {
                // i is wave-uniform: pick the record with selects (indexing
                // the array with the loop counter would put it in scratch)
                BlockMeta bm = bms[0];
#pragma unroll
                for (int q = 1; q < RPW; q++)
                    if (i == q) bm = bms[q];
                const double Vm = LV[r * W + lane];
                const bool pok = pidx[i] >= 0;
                const double Vp = pok ? LV[pidx[i]] : 0.0;
                if constexpr (CHAN) {
                    const CompOut o = RUN::run(d, bm, lane, Vm, pok, Vp, active, step0 + k);
                    LB[r * W + lane] = o.rhs;
                    LD[r * W + lane] = o.diag;
                    if (r == R - 1) droot0 = o.diag;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RMAX - 1; r++) {
            if (r < R - 1) {
                if (wave == r / RPW) {
                    const int i = r % RPW;
                    const int p = par[i];
                    if (p >= 0) {
                        // all four LDS reads issue before the first write
                        // (the compiler cannot tell LB from LD apart)
                        const double dr = LD[r * W + lane], br = LB[r * W + lane];
                        const double dp = LD[p * W + lane], bp = LB[p * W + lane];
                        const double tt = ta[i] / dr;
                        LD[p * W + lane] = dp - tcf[i] * tt;
                        LB[p * W + lane] = bp - br * tt;
                    }
                }
                // a wave's consecutive rows need no barrier between them
                // (its LDS accesses complete in order); the next wave's do
                if (r % RPW == RPW - 1 || r == R - 2) __syncthreads();
            }
        }
#pragma unroll
        for (int r = RMAX - 1; r >= 0; r--) {
            if (r < R) {
                if (wave == r / RPW) {
                    const int i = r % RPW;
                    const int p = par[i];
                    const double dr = LD[r * W + lane];
                    const double xp = p >= 0 ? LB[p * W + lane] : 0.0;
                    double bv = LB[r * W + lane];
                    if (p >= 0) bv -= tcb[i] * xp;
                    const double x = bv / dr;
                    // a one-row solver keeps rhs in results[0] (hines_solve.c:236-238)
                    if (CHAN || !d.single) LB[r * W + lane] = x;
                    if (d.crank) {
                        const double v = LV[r * W + lane];
                        LV[r * W + lane] = x + x - v;
                    } else {
                        LV[r * W + lane] = x;
                    }
                    // the reference never stores its last row's eliminated
                    // diagonal (hines_solve.c:275-276)
                    if (r == R - 1 && is_last) LD[r * W + lane] = droot0;
                }
                if (r % RPW == 0) __syncthreads();
            }
        }
    }
#pragma unroll
    for (int i = 0; i < RPW; i++) {
        const int r = wave * RPW + i;
        if (r < R) {
            const long long b = b0 + r;
            d.V[b * W + lane] = LV[r * W + lane];
            d.R[b * 2 * W + lane] = LB[r * W + lane];
            d.R[b * 2 * W + W + lane] = LD[r * W + lane];
        }
    }
    if (threadIdx.x == 0) d.step[t] = step0 + K;
}

// Batched persistent kernel for trees of up to 64 rows: one workgroup per
// tile, waves split the rows for the channel phase, wave 0 runs the 64
// eliminations (one per lane) with coefficients read from L2.
template <class RUN>
__device__ __forceinline__ void body_batched(const DevModel& d, int K, double* lds)
{
    const int t = blockIdx.x;
    const TileMeta tm = d.tiles[t];
    const int R = tm.rows, b0 = tm.block0;
    const int nw = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & (W - 1);
    double* LV = lds;
    double* LB = lds + R * W;
    double* LD = lds + 2 * R * W;
    for (int r = wave; r < R; r += nw) LV[r * W + lane] = d.V[(long long)(b0 + r) * W + lane];
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
            CompOut o = RUN::run(d, bm, lane, Vm, pok, Vp, lane < tm.lanes, step0 + k);
            LB[r * W + lane] = o.rhs;
            LD[r * W + lane] = o.diag;
        }
        __syncthreads();
        if (wave == 0) {
            // one lane per neuron; the tree coefficients do not depend on the
            // sweep, so they are loaded BCH rows at a time ahead of the rows
            // that use them, and each row's LDS reads issue before its writes
            const long long p0 = (long long)b0 * W + lane;
            const bool is_last = (p0 + (long long)(R - 1) * W) == (long long)d.last_pos;
            const double droot0 = LD[(R - 1) * W + lane];
            for (int r0 = 0; r0 < R - 1; r0 += BCH) {
                int pp[BCH];
                double ca[BCH], cf[BCH];
#pragma unroll
                for (int i = 0; i < BCH; i++) {
                    const int r = r0 + i;
                    pp[i] = -1;
                    ca[i] = cf[i] = 0.0;
                    if (r < R - 1) {
                        const long long q = p0 + (long long)r * W;
                        pp[i] = d.par[q];
                        ca[i] = d.ta[q];
                        cf[i] = d.tcf[q];
                    }
                }
#pragma unroll
                for (int i = 0; i < BCH; i++) {
                    const int r = r0 + i;
                    const int p = pp[i];
                    if (p >= 0) {
                        const double dr = LD[r * W + lane], br = LB[r * W + lane];
                        const double dp = LD[p * W + lane], bp = LB[p * W + lane];
                        const double tt = ca[i] / dr;
                        LD[p * W + lane] = dp - cf[i] * tt;
                        LB[p * W + lane] = bp - br * tt;
                    }
                }
            }
            for (int r1 = R - 1; r1 >= 0; r1 -= BCH) {
                int pp[BCH];
                double cb[BCH];
#pragma unroll
                for (int i = 0; i < BCH; i++) {
                    const int r = r1 - i;
                    pp[i] = -1;
                    cb[i] = 0.0;
                    if (r >= 0) {
                        const long long q = p0 + (long long)r * W;
                        pp[i] = d.par[q];
                        cb[i] = d.tcb[q];
                    }
                }
#pragma unroll
                for (int i = 0; i < BCH; i++) {
                    const int r = r1 - i;
                    if (r >= 0) {
                        const int p = pp[i];
                        const double dr = LD[r * W + lane];
                        const double xp = p >= 0 ? LB[p * W + lane] : 0.0;
                        const double v = LV[r * W + lane];
                        double bv = LB[r * W + lane];
                        if (p >= 0) bv -= cb[i] * xp;
                        const double x = bv / dr;
                        LB[r * W + lane] = x;
                        LV[r * W + lane] = d.crank ? x + x - v : x;
                    }
                }
            }
            if (is_last) LD[(R - 1) * W + lane] = droot0;
        }
        __syncthreads();
    }
    for (int r = wave; r < R; r += nw) {
        const long long b = b0 + r;
        d.V[b * W + lane] = LV[r * W + lane];
        d.R[b * 2 * W + lane] = LB[r * W + lane];
        d.R[b * 2 * W + W + lane] = LD[r * W + lane];
    }
    if (threadIdx.x == 0) d.step[t] = step0 + K;
}

// Orders one wave's LDS writes before its later LDS reads, across lanes.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Packs (lane = compartment): whole neurons of any size, one workgroup per
// pack.  V, rhs and diag of every pack-local position live in LDS for the
// whole launch.  Per step: the waves run the channel programs of the pack's
// blocks, then the Hines elimination runs level by level -- forward, each
// parent pulls its children in ascending row order, which is how the
// reference's funcs[] program builds a row before storing it
// (hines_solve.c:246-273), so every parent sees the same sequence of
// roundings; backward, roots solve and children follow their parents
// (:275-295) -- with a barrier between levels wider than one wave; the
// narrower levels run in wave 0 alone.  CHAN = false is the plain drop-in's
// elimination (rhs/diag read from results, one step).  STAGE = 1 copies the
// pack's level schedule into LDS once per launch; STAGE = 2 copies only its
// indices (level bounds, entries, kid and backward rows: 2-8 bytes per row
// instead of ~36, so more packs fit a CU at once) and leaves the
// coefficients in global memory; without STAGE = 1 each level's entries are
// loaded one level ahead.
template <class RUN, bool CHAN, int STAGE>
__device__ __forceinline__ void body_forest(const DevModel& d, int K, double* lds)
{
    const int pk = blockIdx.x;
    const PackMeta pm = d.packs[pk];
    const int NP = pm.nblk * W;
    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int nw = T >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & (W - 1);
    double* LV = lds;
    double* LB = lds + NP;
    double* LD = lds + 2 * NP;
    char* tail = (char*)(lds + 3 * NP);
    int* LP = (int*)tail;                      // predecessor lp (-1: outside the pack)
    if (CHAN) tail += (size_t)NP * sizeof(int);
    const long long base = (long long)pm.block0 * W;
    const int* lvl = d.lvl + pm.lv0;
    const int* FE = d.fent + 2 * (long long)pm.fe0;
    const unsigned short* KI = d.kidx + pm.ke0;
    const double* KTA = d.kta + pm.ke0;
    const double* KTF = d.ktcf + pm.ke0;
    const int* BKP = d.bkp + pm.be0;
    const double* BTC = d.btcb + pm.be0;
    if (STAGE == 2) {
        int* sLVL = (int*)tail;
        int* sFE = sLVL + (pm.nfl + pm.nbl + 2);
        int* sBKP = sFE + 2 * pm.nfe;
        unsigned short* sKI = (unsigned short*)(sBKP + pm.nbe);
        for (int i = tid; i < pm.nke; i += T) sKI[i] = KI[i];
        for (int i = tid; i < pm.nbe; i += T) sBKP[i] = BKP[i];
        for (int i = tid; i < 2 * pm.nfe; i += T) sFE[i] = FE[i];
        for (int i = tid; i < pm.nfl + pm.nbl + 2; i += T) sLVL[i] = lvl[i];
        lvl = sLVL;
        FE = sFE;
        BKP = sBKP;
        KI = sKI;
    }
    if (STAGE == 1) {
        // doubles first (8-byte aligned), then ints, then shorts
        double* sKTA = (double*)tail;
        double* sKTF = sKTA + pm.nke;
        double* sBTC = sKTF + pm.nke;
        int* sLVL = (int*)(sBTC + pm.nbe);
        int* sFE = sLVL + (pm.nfl + pm.nbl + 2);
        int* sBKP = sFE + 2 * pm.nfe;
        unsigned short* sKI = (unsigned short*)(sBKP + pm.nbe);
        for (int i = tid; i < pm.nke; i += T) {
            sKTA[i] = KTA[i];
            sKTF[i] = KTF[i];
            sKI[i] = KI[i];
        }
        for (int i = tid; i < pm.nbe; i += T) {
            sBTC[i] = BTC[i];
            sBKP[i] = BKP[i];
        }
        for (int i = tid; i < 2 * pm.nfe; i += T) sFE[i] = FE[i];
        for (int i = tid; i < pm.nfl + pm.nbl + 2; i += T) sLVL[i] = lvl[i];
        KTA = sKTA;
        KTF = sKTF;
        BTC = sBTC;
        lvl = sLVL;
        FE = sFE;
        BKP = sBKP;
        KI = sKI;
    }
    // several positions' loads in flight before their LDS stores wait
#pragma unroll 4
    for (int i = tid; i < NP; i += T) {
        LV[i] = d.V[base + i];
        if (CHAN) {
            const int pp = d.pred[base + i];
            LP[i] = (pp >= base && pp < base + NP) ? (int)(pp - base) : -1;
        } else {
            const long long r = (long long)(pm.block0 + (i >> 6)) * 2 * W + (i & (W - 1));
            LB[i] = d.R[r];
            LD[i] = d.R[r + W];
        }
    }
    const int step0 = d.pstep[pk];
    int kstep = 0;     // the step of the launch being run (for the adjoint's recording)
    // The reference never stores its last row's eliminated diagonal
    // (hines_solve.c:275-276): the thread that produces that row's diag keeps
    // a copy and puts it back after the last step.
    double droot0 = 0.0;
    bool own_last = false;
    if (!CHAN && pm.last_lp >= 0 && pm.last_lp % T == tid) {
        own_last = true;
        droot0 = LD[pm.last_lp];
    }
    __syncthreads();
    // Forward entry: a parent folds its kids in the reference's order.  The
    // load part reads only what no other entry of the sweep writes (the
    // schedule, the parent's own rhs/diag), so it is issued one level ahead;
    // the run part reads the kids' values the previous level produced.
    struct FwdE {
        int p, kb, nk;
        int c[3];
        double ta[3], tf[3];
        double dp, bp;
    };
    auto fwd_load = [&](int e, FwdE& x) {
        const int f0 = FE[2 * e];
        x.kb = FE[2 * e + 1];
        x.p = f0 & 0xffff;
        x.nk = f0 >> 16;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int j = q < x.nk ? x.kb + q : x.kb;
            x.c[q] = KI[j];
            x.ta[q] = KTA[j];
            x.tf[q] = KTF[j];
        }
        x.dp = LD[x.p];
        x.bp = LB[x.p];
    };
    // The kids' values (up to three), produced by the previous level.
    struct FwdK {
        double kd[3], kb[3];
    };
    auto fwd_kids = [&](const FwdE& x, FwdK& k) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            k.kd[q] = LD[x.c[q]];
            k.kb[q] = LB[x.c[q]];
        }
    };
    auto fwd_run = [&](const FwdE& x, const FwdK& k) {
        double dp = x.dp, bp = x.bp;
        if (x.nk <= 3) {
            // up to three kids: the independent divides first, then the
            // updates in the reference's order
            double tt[3];
#pragma unroll
            for (int q = 0; q < 3; q++) tt[q] = x.ta[q] / k.kd[q];
#pragma unroll
            for (int q = 0; q < 3; q++)
                if (q < x.nk) {
                    dp -= x.tf[q] * tt[q];
                    bp -= k.kb[q] * tt[q];
                }
        } else {
            for (int j = x.kb; j < x.kb + x.nk; j++) {
                const int c = KI[j];
                const double tt = KTA[j] / LD[c];
                dp -= KTF[j] * tt;
                bp -= LB[c] * tt;
            }
        }
        LD[x.p] = dp;
        LB[x.p] = bp;
    };
    // Backward entry: a row's solution from its parent's.  Only the parent's
    // solution comes from the previous level; the rest is loaded ahead.
    struct BwdE {
        int c, p;
        double tcb, bc, dc, v;
    };
    auto bwd_load = [&](int e, BwdE& x) {
        const int kp = BKP[e];
        x.c = kp & 0xffff;
        x.p = (kp >> 16) - 1;
        x.tcb = BTC[e];
        x.bc = LB[x.c];
        x.dc = LD[x.c];
        x.v = d.crank ? LV[x.c] : 0.0;
    };
    auto bwd_par = [&](const BwdE& x) { return x.p >= 0 ? LB[x.p] : 0.0; };
    auto bwd_run = [&](const BwdE& x, double xp) {
        double bv = x.bc;
        if (x.p >= 0) bv -= x.tcb * xp;
        const double xs = bv / x.dc;
        // the adjoint's recording: eliminated rhs and diag, solution
        if (CHAN && d.recB) {
            const long long o = (long long)kstep * d.rec_npos + base + x.c;
            d.recB[o] = x.bc;
            d.recD[o] = x.dc;
            d.recX[o] = xs;
        }
        // a one-row solver keeps rhs in results[0] (hines_solve.c:236-238)
        if (!d.single) LB[x.c] = xs;
        LV[x.c] = d.crank ? xs + xs - x.v : xs;
    };
    // Levels [l0, l1) of a sweep (level l: entries lv[l] .. lv[l+1]-1).
    // ALL: every thread, a workgroup barrier after each level.  Otherwise
    // wave 0 alone, whose LDS accesses complete in order, so its levels need
    // no barrier.  With the coefficients in global memory (STAGE != 1) each
    // thread's first entry of the next level is loaded before the current
    // level runs, taking the schedule's memory latency off the level chain;
    // with the schedule in LDS that would only queue the next level's LDS
    // reads ahead of this level's (one wave's LDS reads complete in order):
    // measured slower (12.4 -> 13.9 us per step for 512 packs of 128).
    auto fwd_levels = [&](int l0, int l1, bool all) {
        const int t0 = all ? tid : lane, st = all ? T : W;
        FwdE nx;
        bool vn = false;
        if (STAGE != 1 && l0 < l1) {
            const int e = lvl[l0] + t0;
            vn = e < lvl[l0 + 1];
            if (vn) fwd_load(e, nx);
        }
        for (int l = l0; l < l1; l++) {
            int e1 = lvl[l] + t0;
            if (STAGE != 1) {
                const FwdE cu = nx;
                const bool vc = vn;
                if (l + 1 < l1) {
                    const int e = lvl[l + 1] + t0;
                    vn = e < lvl[l + 2];
                    if (vn) fwd_load(e, nx);
                }
                if (vc) {
                    FwdK k;
                    fwd_kids(cu, k);
                    fwd_run(cu, k);
                }
                e1 += st;
            }
            for (int e = e1; e < lvl[l + 1]; e += st) {
                FwdE x;
                FwdK k;
                fwd_load(e, x);
                fwd_kids(x, k);
                fwd_run(x, k);
            }
            if (all)
                __syncthreads();
            else
                wave_lds_sync();
        }
    };
    auto bwd_levels = [&](const int* lv, int l0, int l1, bool all) {
        const int t0 = all ? tid : lane, st = all ? T : W;
        BwdE nx;
        bool vn = false;
        if (STAGE != 1 && l0 < l1) {
            const int e = lv[l0] + t0;
            vn = e < lv[l0 + 1];
            if (vn) bwd_load(e, nx);
        }
        for (int l = l0; l < l1; l++) {
            int e1 = lv[l] + t0;
            if (STAGE != 1) {
                const BwdE cu = nx;
                const bool vc = vn;
                if (l + 1 < l1) {
                    const int e = lv[l + 1] + t0;
                    vn = e < lv[l + 2];
                    if (vn) bwd_load(e, nx);
                }
                if (vc) bwd_run(cu, bwd_par(cu));
                e1 += st;
            }
            for (int e = e1; e < lv[l + 1]; e += st) {
                BwdE x;
                bwd_load(e, x);
                bwd_run(x, bwd_par(x));
            }
            if (all)
                __syncthreads();
            else
                wave_lds_sync();
        }
    };
    // Levels [l0, l1) of at most one wave's width, run by wave 0 (one entry
    // per lane at most), when there are fewer than 64 of them: the levels'
    // bounds are held one per lane and read with readlane, and each lane's
    // entry of the next level -- its schedule indices -- is read while the
    // running level waits for its own reads, into the other of two records
    // (so nothing carried to the next level is copied, which would wait).
    // A level then waits for one round of LDS reads (kids' values or the
    // parent's solution, with the coefficients and the entry's own values)
    // instead of four (bounds, entry, kid indices, kids' values).
    struct FwdI {
        int p, kb, nk, c0, c1, c2;
        bool v;
    };
    auto fwd_narrow = [&](int l0, int l1) {
        const int nl = l1 - l0;
        const int bnd = lane <= nl ? lvl[l0 + lane] : 0;
        auto fidx = [&](int j, FwdI& x) {
            const int e = __builtin_amdgcn_readlane(bnd, j) + lane;
            x.v = e < __builtin_amdgcn_readlane(bnd, j + 1);
            if (x.v) {
                const int f0 = FE[2 * e];
                x.kb = FE[2 * e + 1];
                x.p = f0 & 0xffff;
                x.nk = f0 >> 16;
                x.c0 = KI[x.kb];
                x.c1 = KI[x.nk > 1 ? x.kb + 1 : x.kb];
                x.c2 = KI[x.nk > 2 ? x.kb + 2 : x.kb];
            }
        };
        auto level = [&](const FwdI& a, int jn, FwdI& b) {
            FwdE x;
            FwdK k;
            if (a.v) {
                x.p = a.p;
                x.kb = a.kb;
                x.nk = a.nk;
                x.c[0] = a.c0;
                x.c[1] = a.c1;
                x.c[2] = a.c2;
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const int j = q < x.nk ? x.kb + q : x.kb;
                    x.ta[q] = KTA[j];
                    x.tf[q] = KTF[j];
                }
                x.dp = LD[x.p];
                x.bp = LB[x.p];
                fwd_kids(x, k);
            }
            if (jn < nl) fidx(jn, b);
            if (a.v) fwd_run(x, k);
            wave_lds_sync();
        };
        FwdI A, B;
        fidx(0, A);
        for (int j = 0; j < nl; j += 2) {
            level(A, j + 1, B);
            if (j + 1 < nl) level(B, j + 2, A);
        }
    };
    struct BwdI {
        int c, p, e;
        bool v;
    };
    auto bwd_narrow = [&](const int* lv, int l0, int l1) {
        const int nl = l1 - l0;
        const int bnd = lane <= nl ? lv[l0 + lane] : 0;
        auto bidx = [&](int j, BwdI& x) {
            x.e = __builtin_amdgcn_readlane(bnd, j) + lane;
            x.v = x.e < __builtin_amdgcn_readlane(bnd, j + 1);
            if (x.v) {
                const int kp = BKP[x.e];
                x.c = kp & 0xffff;
                x.p = (kp >> 16) - 1;
            }
        };
        auto level = [&](const BwdI& a, int jn, BwdI& b) {
            BwdE x;
            double xp = 0.0;
            if (a.v) {
                x.c = a.c;
                x.p = a.p;
                x.tcb = BTC[a.e];
                x.bc = LB[x.c];
                x.dc = LD[x.c];
                x.v = d.crank ? LV[x.c] : 0.0;
                xp = bwd_par(x);
            }
            if (jn < nl) bidx(jn, b);
            if (a.v) bwd_run(x, xp);
            wave_lds_sync();
        };
        BwdI A, B;
        bidx(0, A);
        for (int j = 0; j < nl; j += 2) {
            level(A, j + 1, B);
            if (j + 1 < nl) level(B, j + 2, A);
        }
    };
    // Levels of at most one wave's width at the root end of the forward
    // sweep and at either end of the backward sweep run in wave 0 alone:
    // one wave's LDS accesses complete in order, so those levels need no
    // workgroup barrier between them.  Wider levels use every wave.
    const int* blv = lvl + pm.nfl + 1;
    int fwide = pm.nfl;
    while (fwide > 0 && lvl[fwide] - lvl[fwide - 1] <= W) fwide--;
    int bhead = 0;
    while (bhead < pm.nbl && blv[bhead + 1] - blv[bhead] <= W) bhead++;
    int btail = pm.nbl;
    while (btail > bhead && blv[btail] - blv[btail - 1] <= W) btail--;
    fwide = __builtin_amdgcn_readfirstlane(fwide);
    bhead = __builtin_amdgcn_readfirstlane(bhead);
    btail = __builtin_amdgcn_readfirstlane(btail);
    for (int k = 0; k < K; k++) {
        kstep = k;
        if (CHAN) {
            // the adjoint's recomputation keeps the state each step starts from
            if (d.recV)
                for (int i = tid; i < NP; i += T) d.recV[k * d.rec_npos + base + i] = LV[i];
            for (int blk = wave; blk < pm.nblk; blk += nw) {
                const BlockMeta bm = d.blocks[pm.block0 + blk];
                const int i = blk * W + lane;
                if (d.recS) {
                    const int ns = d.prog[bm.prog];
                    const double* src = d.S + bm.dslot + lane;
                    double* dst = d.recS + k * d.rec_nS + bm.dslot + lane;
                    for (int c = 0; c < ns; c++) dst[c * W] = src[c * W];
                }
                const double Vm = LV[i];
                const int pi = LP[i];
                const bool pok = pi >= 0;
                const double Vp = pok ? LV[pi] : 0.0;
                CompOut o = RUN::run(d, bm, lane, Vm, pok, Vp, lane < bm.lanes, step0 + k);
                LB[i] = o.rhs;
                LD[i] = o.diag;
                if (i == pm.last_lp) {
                    own_last = true;
                    droot0 = o.diag;
                }
            }
            __syncthreads();
        }
        // forward levels that need every wave, then the narrow tail of the
        // forward sweep and the narrow head of the backward sweep in wave 0
        fwd_levels(0, fwide, true);
        if (fwide < pm.nfl || bhead > 0) {
            if (wave == 0) {
                if (pm.nfl - fwide < W) fwd_narrow(fwide, pm.nfl);
                else fwd_levels(fwide, pm.nfl, false);
                if (bhead < W) bwd_narrow(blv, 0, bhead);
                else bwd_levels(blv, 0, bhead, false);
            }
            __syncthreads();
        }
        bwd_levels(blv, bhead, btail, true);
        if (btail < pm.nbl) {
            if (wave == 0) {
                if (pm.nbl - btail < W) bwd_narrow(blv, btail, pm.nbl);
                else bwd_levels(blv, btail, pm.nbl, false);
            }
            __syncthreads();
        }
        // the adjoint's forward run keeps the observed voltages of each step
        if (CHAN && d.trace)
            for (int i = tid; i < NP; i += T) {
                const int j = d.obs_idx[base + i];
                if (j >= 0) d.trace[(long long)k * d.nobs + j] = LV[i];
            }
    }
    if (own_last) LD[pm.last_lp] = droot0;
    __syncthreads();
    for (int i = tid; i < NP; i += T) {
        d.V[base + i] = LV[i];
        const long long r = (long long)(pm.block0 + (i >> 6)) * 2 * W + (i & (W - 1));
        d.R[r] = LB[i];
        d.R[r + W] = LD[i];
    }
    if (tid == 0) d.pstep[pk] = step0 + K;
}

// The spike pass of step `step` for one source (hines_chip.c:266-276
// SPIKE_OP, which reads only Vm -- the last step's solution -- its threshold
// and its counter): counter - 1; when v > thresh with the counter spent,
// record the spike, reload the counter and add each of the source's
// connections (its in-connection slot) to the target synchan's arrivals of
// that step (the buffers of step's parity).  Every lane of the wave calls it,
// live or not: the wave scatters its spiking lanes' connections together, so
// a wave's atomics go to consecutive arrivals.
__device__ __forceinline__ void spike_src_pass(const DevModel& d, bool live, const SpikeSrc& s, double v, int step)
{
    bool fired = false;
    if (live) {
        int cnt = d.I[s.cpos] - 1;
        if (v > d.S[s.spos] && cnt <= 0) {
            const int q = atomicAdd(d.nspk, 1);
            if (q < d.spk_cap) {
                d.spk[2 * q] = s.opidx;
                d.spk[2 * q + 1] = step;
            }
            cnt = d.I[s.rpos];
            fired = true;
        }
        d.I[s.cpos] = cnt;
    }
    const int lane = threadIdx.x & (W - 1);
    unsigned long long m = __ballot(fired);
    while (m) {
        const int l = __ffsll((long long)m) - 1;
        m &= m - 1;
        // every lane scatters lane l's connections, into the buffers of
        // lane l's step (a lane past the last source holds no step)
        const int o0 = __shfl(s.out0, l), o1 = __shfl(s.out1, l);
        const int par = __shfl(step, l) & 1;
        int* __restrict__ narr = d.narr + par * d.ev_nsyn;
        int* __restrict__ arr = d.arr + par * d.ev_nconn;
        for (int o = o0 + lane; o < o1; o += W) {
            const int sid = d.out_sid[o];
            const int k = atomicAdd(narr + sid, 1);
            arr[d.in_off[sid] + k] = d.out_slot[o];
        }
    }
}

// The clock of the spike pass of step `step`: SimulationTime() advances by
// dt (sim/sim_step.c:112); the synchans of that step read tstep[step & 1].
// One thread per step calls it.
__device__ __forceinline__ void ev_clock(const DevModel& d, int step)
{
    EvCtl* c = d.evctl;
    const double t = c->tnext;
    c->tlast = c->tnow;
    c->tnow = t;
    c->tstep[step & 1] = t;
    c->tnext = t + d.dt;
}

// The spike pass of step `step` for the sources of block b, run by the wave
// that has just solved the block: V is its lane's new voltage, which is all
// the next step's SPIKE_OPs read of the step, so the next step needs no
// launch of its own for them.  Block 0's wave advances the clock.  The
// arrivals go to the buffers of step's parity, which no synchan of the
// running step reads.
__device__ __forceinline__ void spike_lookahead(const DevModel& d, int b, int lane, double V, int step)
{
    if (b == 0 && lane == 0) ev_clock(d, step);
    const int s0 = d.src_boff[b], s1 = d.src_boff[b + 1];
    for (int i0 = s0; i0 < s1; i0 += W) {
        const int i = i0 + lane;
        const bool live = i < s1;
        SpikeSrc q{};
        if (live) q = d.src[i];
        const double v = __shfl(V, q.vpos & (W - 1), W);
        spike_src_pass(d, live, q, v, step);
    }
}

// Isopotential cells (every tree one compartment): one lane per cell, the
// voltage stays in a register for K steps.  V = rhs/diag is the reference's
// single-row solve (hines_solve.c:236-238; CN :158-162).
template <class RUN>
__device__ __forceinline__ void body_iso(const DevModel& d, int K)
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
        o = RUN::run(d, bm, lane, V, pok, Vp, lane < tm.lanes, step0 + k);
        x = o.rhs / o.diag;
        if (d.crank) V = x + x - V;
        else V = x;
    }
    d.V[pos] = V;
    // a one-row solver keeps rhs in results[0] (hines_solve.c:236-238)
    d.R[(long long)b * 2 * W + lane] = d.single ? o.rhs : x;
    d.R[(long long)b * 2 * W + W + lane] = o.diag;
    if (lane == 0) d.step[t] = step0 + K;
    if (d.look) spike_lookahead(d, b, lane, V, step0 + K);
}

}  // namespace hsh
