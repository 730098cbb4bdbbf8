/*
 * hsolve_tl.c -- tangent-linear (forward-mode derivative) restatement of the
 * hsolve chanmode-4 step, for checking gradients.
 *
 * TEST INFRASTRUCTURE ONLY (see hsolve_oracle.h).
 *
 * Each step runs the reference step exactly as hsolve_oracle.c does
 * (do_chip_hh4_update, genesis/src/hines/hines_chip.c:157-600, and
 * do_euler_hsolve / do_crank_hsolve, hines_solve.c:140-296, walking ops[],
 * chip[], funcs[] and ravals[] in the reference order) and carries beside
 * every value its derivative along one direction: a tangent of the initial
 * vm[], of any chip[] entry (gate states, gmax, Ek, Em/Rm, inject, tbyc,
 * diagterm) and of ravals[].  Table interpolation is piecewise linear in Vm,
 * so d(table)/dVm = (next row - row) * invdx inside the table and 0 where Vm
 * is clamped.  Supported ops: the tabchannel subset (NEWVOLT, CHAN, CHAN_EK,
 * IPOL1V with powers 1..6 and instant gates, ADD_CURR) and SPIKE (no
 * derivative); anything else is rejected at create time.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hsolve_oracle.h"

#define LCOMPT_OP 102
#define SPIKE_OP 200
#define CHAN_OP 3000
#define CHAN_EK_OP 3001
#define ADD_CURR_OP 3200
#define IPOL1V_OP 4001
#define NEWVOLT_OP 5100
#define FORWARD_ELIM 0
#define BACKWARD_ELIM 0
#define SET_DIAG 1
#define SKIP_DIAG 2
#define CALC_RESULTS 6
#define CRANK_INT 11

struct ohs_tl {
    ohs_arrays a;
    int32_t *ops;
    double *chip, *dchip, *vm, *dvm, *ravals, *dravals;
    double *res, *dres;
};

static void *dupz(const void *p, size_t n)
{
    void *q = calloc(n ? n : 1, 1);
    if (p && n) memcpy(q, p, n);
    return q;
}

/* Check that ops[] holds only the supported subset. */
static int supported(const ohs_arrays *a)
{
    size_t p = 0;
    while (p < (size_t)a->nops && a->ops[p] != LCOMPT_OP) {
        p++;     /* COMPT_OP / FCOMPT_OP */
        while (p < (size_t)a->nops && a->ops[p] > LCOMPT_OP) {
            int op = a->ops[p++];
            if (op == NEWVOLT_OP || op == ADD_CURR_OP || op == CHAN_OP || op == CHAN_EK_OP) continue;
            if (op == IPOL1V_OP) { p += 2; continue; }
            if (op == SPIKE_OP) { p += 2; continue; }
            return 0;
        }
    }
    /* funcs[]: forward part up to FINISH, backward part up to FINISH; the
     * sibling ops of symcompartments are not supported */
    int32_t f = 0;
    for (int pass = 0; pass < 2 && a->ncompts > 1; pass++) {
        for (;;) {
            if (f >= a->nfuncs) return 0;
            int op = a->funcs[f++];
            if (op == FORWARD_ELIM) f++;
            else if (op == SET_DIAG || op == SKIP_DIAG || op == CALC_RESULTS) continue;
            else if (op == 7) break;
            else return 0;
        }
    }
    return 1;
}

ohs_tl *ohs_tl_create(const ohs_arrays *a)
{
    if (!supported(a)) return NULL;
    ohs_tl *t = (ohs_tl *)calloc(1, sizeof(ohs_tl));
    t->a = *a;
    t->ops = (int32_t *)dupz(a->ops, sizeof(int32_t) * a->nops);
    t->chip = (double *)dupz(a->chip, sizeof(double) * a->nchips);
    t->dchip = (double *)calloc(a->nchips + 1, sizeof(double));
    t->vm = (double *)dupz(a->vm, sizeof(double) * a->ncompts);
    t->dvm = (double *)calloc(a->ncompts + 1, sizeof(double));
    t->ravals = (double *)dupz(a->ravals, sizeof(double) * a->nravals);
    t->dravals = (double *)calloc(a->nravals + 1, sizeof(double));
    t->res = (double *)calloc(2 * (size_t)a->ncompts + 2, sizeof(double));
    t->dres = (double *)calloc(2 * (size_t)a->ncompts + 2, sizeof(double));
    return t;
}

void ohs_tl_destroy(ohs_tl *t)
{
    if (!t) return;
    free(t->ops); free(t->chip); free(t->dchip); free(t->vm); free(t->dvm);
    free(t->ravals); free(t->dravals); free(t->res); free(t->dres);
    free(t);
}

void ohs_tl_set_tangent(ohs_tl *t, const double *dvm, const double *dchip, const double *dravals)
{
    if (dvm) memcpy(t->dvm, dvm, sizeof(double) * t->a.ncompts);
    else memset(t->dvm, 0, sizeof(double) * t->a.ncompts);
    if (dchip) memcpy(t->dchip, dchip, sizeof(double) * t->a.nchips);
    else memset(t->dchip, 0, sizeof(double) * t->a.nchips);
    if (dravals) memcpy(t->dravals, dravals, sizeof(double) * t->a.nravals);
    else memset(t->dravals, 0, sizeof(double) * t->a.nravals);
}

/* Gk *= X^k with the reference's multiplication sequence
 * (hines_chip.c:490-509), and its tangent. */
static void power_tl(double *Gk, double *dGk, double X, double dX, int k)
{
    double Xp = 1.0;
    for (int i = 0; i < k; i++) Xp *= X;
    double dXp = k * dX;
    for (int i = 0; i < k - 1; i++) dXp *= X;
    *dGk = *dGk * Xp + *Gk * dXp;
    double G = *Gk;
    if (k == 1) {
        G *= X;
    } else if (k == 2) {
        G *= X * X;
    } else if (k == 3) {
        G *= X * X * X;
    } else if (k == 4) {
        double Y = X * X;
        G *= Y * Y;
    } else if (k == 5) {
        double Y = X;
        G *= Y;
        Y *= Y;
        Y *= Y;
        G *= Y;
    } else {
        double Y = X * X;
        G *= Y;
        Y *= Y;
        G *= Y;
    }
    *Gk = G;
}

static void chip_update_tl(ohs_tl *t)
{
    const ohs_arrays *A = &t->a;
    int32_t *ops = t->ops;
    double *chip = t->chip, *dchip = t->dchip;
    const int ncols = A->ncols, xdivs = A->xdivs;
    const double xmin = A->xmin, invdx = A->invdx;
    size_t p = 0, c = 0, r = 0, vi = 0;
    int filo = 0, inside = 0;
    double xlo = 0.0, xhi = 0.0, vipol = 0.0, dvipol = 0.0;
    const double *vtab = NULL;
    double Gk = 0.0, dGk = 0.0, Ek = 0.0, dEk = 0.0;
    if (!A->tablist) { xlo = -1e20; xhi = 1e20; }
    while (ops[p++] != LCOMPT_OP) {
        double Vm = t->vm[vi], dVm = t->dvm[vi];
        vi++;
        double sumg = 0.0, dsumg = 0.0;
        double ichan = chip[c] + chip[c + 1], dichan = dchip[c] + dchip[c + 1];
        c += 2;
        while (ops[p] > LCOMPT_OP) {
            int op = ops[p++];
            if (op == NEWVOLT_OP) {
                double raw = (Vm - xmin) * invdx;
                vipol = raw;
                if ((xlo > Vm) || (Vm >= xhi)) {
                    if (vipol < 0.0) vipol = 0.0;
                    else if (vipol > xdivs) vipol = xdivs;
                    filo = (int)vipol;
                    vipol = vipol - (double)filo;
                    xlo = A->xvals[filo];
                    xhi = A->xvals[filo + 1];
                    vtab = A->tablist + (size_t)filo * ncols;
                } else {
                    vipol = vipol - (double)filo;
                }
                inside = (raw > 0.0 && raw < (double)xdivs);
                dvipol = inside ? dVm * invdx : 0.0;
            } else if (op == CHAN_EK_OP) {
                Ek = chip[c];
                dEk = dchip[c];
                c++;
                Gk = chip[c];
                dGk = dchip[c];
                c++;
            } else if (op == CHAN_OP) {
                Gk = chip[c];
                dGk = dchip[c];
                c++;
            } else if (op == IPOL1V_OP) {
                const double *tab = vtab + ops[p++];
                int k = ops[p++];
                double B = tab[0] + vipol * (tab[ncols] - tab[0]);
                double dB = dvipol * (tab[ncols] - tab[0]);
                double Aa = tab[1] + vipol * (tab[ncols + 1] - tab[1]);
                double dA = dvipol * (tab[ncols + 1] - tab[1]);
                double X, dX;
                if (k < 0) {
                    X = Aa / B;
                    dX = (dA - X * dB) / B;
                    k = -k;
                } else {
                    double X0 = chip[c], dX0 = dchip[c];
                    X = (X0 * (2.0 - B) + Aa) / B;
                    dX = (dX0 * (2.0 - B) - X0 * dB + dA - X * dB) / B;
                }
                chip[c] = X;
                dchip[c] = dX;
                c++;
                power_tl(&Gk, &dGk, X, dX, k);
            } else if (op == ADD_CURR_OP) {
                sumg += Gk;
                dsumg += dGk;
                ichan += Ek * Gk;
                dichan += dEk * Gk + Ek * dGk;
            } else if (op == SPIKE_OP) {
                ops[p] -= 1;
                if (Vm > chip[c++]) {
                    if (ops[p] <= 0) ops[p] = ops[p + 1];
                }
                p += 2;
            }
        }
        double tbyc = chip[c], dtbyc = dchip[c];
        t->res[r] = Vm + ichan * tbyc;
        t->dres[r] = dVm + dichan * tbyc + ichan * dtbyc;
        r++;
        t->res[r] = sumg * tbyc + chip[c + 1];
        t->dres[r] = dsumg * tbyc + sumg * dtbyc + dchip[c + 1];
        r++;
        c += 2;
    }
}

/* do_euler_hsolve / do_crank_hsolve (hines_solve.c:140-296) with tangents;
 * returns the solved x (before the Crank-Nicolson 2x - V) in res. */
static void solve_tl(ohs_tl *t)
{
    const int32_t *funcs = t->a.funcs;
    double *res = t->res, *dres = t->dres;
    const double *ra = t->ravals, *dra = t->dravals;
    const int n = t->a.ncompts;
    const int crank = (t->a.method == CRANK_INT);
    double *vm = t->vm, *dvm = t->dvm;
#define OUT(i, x, dx)                                       \
    do {                                                    \
        if (crank) {                                        \
            double v_ = vm[i], dv_ = dvm[i];                \
            vm[i] = (x) + (x) - v_;                         \
            dvm[i] = (dx) + (dx) - dv_;                     \
        } else {                                            \
            vm[i] = (x);                                    \
            dvm[i] = (dx);                                  \
        }                                                   \
    } while (0)
    if (n == 1) {
        double x = res[0] / res[1];
        double dx = (dres[0] - x * dres[1]) / res[1];
        OUT(0, x, dx);
        return;
    }
    size_t f = 0, a = 0, row = 2;
    double temp = 0.0, dtemp = 0.0;
    double rv = res[row], drv = dres[row], dv = res[row + 1], ddv = dres[row + 1];
    int op = funcs[f++];
    for (;;) {
        if (op == FORWARD_ELIM) {
            const int kr = funcs[f++];
            double dk = res[kr + 1], ddk = dres[kr + 1];
            temp = ra[a] / dk;
            dtemp = (dra[a] - temp * ddk) / dk;
            a++;
            dv -= ra[a] * temp;
            ddv -= dra[a] * temp + ra[a] * dtemp;
            a++;
            drv -= dres[kr] * temp + res[kr] * dtemp;
            rv -= res[kr] * temp;
        } else if (op == SET_DIAG || op == SKIP_DIAG) {
            res[row] = rv;
            dres[row] = drv;
            res[row + 1] = dv;
            dres[row + 1] = ddv;
            row += (op == SET_DIAG) ? 2 : 4;
            rv = res[row];
            drv = dres[row];
            dv = res[row + 1];
            ddv = dres[row + 1];
        } else {
            break;
        }
        op = funcs[f++];
    }
    int vi = n - 1;
    double x = rv / dv;
    double dx = (drv - x * ddv) / dv;
    res[row] = x;
    dres[row] = dx;
    OUT(vi, x, dx);
    row -= 2;
    rv = res[row];
    drv = dres[row];
    for (;;) {
        op = funcs[f++];
        if (op == BACKWARD_ELIM) {
            const int pr = funcs[f++];
            drv -= dra[a] * res[pr] + ra[a] * dres[pr];
            rv -= ra[a] * res[pr];
            a++;
        } else if (op == CALC_RESULTS) {
            vi--;
            x = rv / res[row + 1];
            dx = (drv - x * dres[row + 1]) / res[row + 1];
            res[row] = x;
            dres[row] = dx;
            OUT(vi, x, dx);
            row -= 2;
            rv = res[row];
            drv = dres[row];
        } else {
            break;
        }
    }
#undef OUT
}

int32_t ohs_tl_step(ohs_tl *t, int32_t nsteps, const int32_t *obs, int32_t nobs, double *trace,
                    double *dtrace)
{
    for (int32_t s = 0; s < nsteps; s++) {
        chip_update_tl(t);
        solve_tl(t);
        for (int32_t j = 0; j < nobs; j++) {
            if (trace) trace[(size_t)s * nobs + j] = t->vm[obs[j]];
            if (dtrace) dtrace[(size_t)s * nobs + j] = t->dvm[obs[j]];
        }
    }
    return 0;
}

void ohs_tl_get(ohs_tl *t, double *vm, double *dvm, double *chip, double *dchip)
{
    if (vm) memcpy(vm, t->vm, sizeof(double) * t->a.ncompts);
    if (dvm) memcpy(dvm, t->dvm, sizeof(double) * t->a.ncompts);
    if (chip) memcpy(chip, t->chip, sizeof(double) * t->a.nchips);
    if (dchip) memcpy(dchip, t->dchip, sizeof(double) * t->a.nchips);
}
