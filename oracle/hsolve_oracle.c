/*
 * hsolve_oracle.c -- CPU parity oracle for the hsolve chanmode-4/5 step.
 *
 * TEST INFRASTRUCTURE ONLY (see hsolve_oracle.h).  Restates, in the same
 * array-of-structures order and with the same floating-point expression
 * shapes, the reference CPU code:
 *   - channel-opcode interpreter: genesis/src/hines/hines_chip.c:78-602
 *     (built as do_chip_hh4_update via hines_4chip.c:19-24);
 *   - the concentration program: genesis/src/hines/hines_conc_chip.c:59-222
 *     (do_h4_conc_chip_update: shells' results and diagonals, pumps, pool
 *     buffers, pools, Ca_concen);
 *   - the diffusion systems' solve: genesis/src/hines/hines_conc.c:39-396
 *     (do_h_conc_solve: tridiagonal for shells alone, banded with buffers);
 *   - Euler / Crank-Nicolson elimination: genesis/src/hines/hines_solve.c
 *     :223-296 and :140-218;
 *   - step phase order: genesis/src/hines/hines.c:222-270.
 * Interpreter registers persist across compartments exactly as in the
 * reference (Ek, Gk, Ik, GHK, k, concin/fluxindex, the table-row caches).
 */
#include "hsolve_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* opcodes: genesis/src/hines/hines_defs.h:178-222 */
#define COMPT_OP 100
#define FCOMPT_OP 101
#define LCOMPT_OP 102
#define SPIKE_OP 200
#define ELMMPUMP_OP 500
#define CONC_VAL_OP 1000
#define NERNST_OP 2000
#define GHK_OP 2010
#define CHAN_OP 3000
#define CHAN_EK_OP 3001
#define SAME_GHK_OP 3002
#define DO_TABGHK_OP 3003
#define ADD_CURR_OP 3200
#define FLUX_OP 3300
#define SCALE_FLUX_OP 3301
#define STORE_ALLI_OP 3501
#define STORE_ALLC_OP 3502
#define IPOL1V_OP 4001
#define IPOL1C_OP 4002
#define IPOL2V_OP 4003
#define IPOL2C_OP 4004
#define SYN2_OP 4100
#define SYN3_OP 4101
#define MGBLOCK_OP 4200
#define NEWVOLT_OP 5100
#define NEWCONC1_OP 5110
#define NEWCONC2_OP 5111
#define DOMAIN_OP 5112
/* concops[] codes: hines_defs.h:209-222 */
#define CONC_OP 10100
#define FCONC_OP 10101
#define LCONC_OP 10102
#define FLUXCONC_OP 10103
#define CACONC_OP 10200
#define LCACONC_OP 10202
#define TAUPUMP_OP 11000
#define CTAUPUMP_OP 11001
#define MMPUMP_OP 11010
#define HILLPUMP_OP 11011
#define PHILLPUMP_OP 11012
#define BUFFER_OP 12000
#define POOL_OP 12100
/* funcs codes: hines_defs.h:119-128 */
#define FORWARD_ELIM 0
#define BACKWARD_ELIM 0
#define SET_DIAG 1
#define SKIP_DIAG 2
#define FASTSIBARRAY_ELIM 3
#define COPY_ARRAY 4
#define SIBARRAY_ELIM 5
#define CALC_RESULTS 6
#define FINISH 7

#define OHS_TINY 1.0e-6   /* hines_defs.h:114 */
#define CRANK_INT 11

/* One synchan's pending synaptic events (newconn/synchan.c:161-316): the
 * list in order, each event's time field (steps after the event before it;
 * for the head, steps counted from time_last_event) and weight/dt of its
 * synapse; time_last_event; the ops index of the synchan's SYN2_OP. */
typedef struct {
    int32_t time;
    double wdt;
} ohs_ev;

typedef struct {
    ohs_ev *ev;
    int32_t n, cap;
    double tle;
    int32_t synop;
} ohs_pend;

struct ohs {
    ohs_arrays a;
    int32_t *ops;
    double *chip, *vm, *results, *givals, *conc, *flux, *codiag, *itotal;
    double *coresult;          /* per concentration: the concops result (hines_conc_chip.c:88) */
    double *dscr_cv, *dscr_rs; /* the solve's elimination scratch (covals, cores) */
    double *results_base;
    int32_t bad_funcs;         /* a funcs[] program that left the row range (not run further) */
    double *ravals;
    int32_t *spike_op, *spike_step;
    int32_t nspk, capspk;
    int32_t step;
    ohs_syn_fn syn;
    ohs_rand_fn rnd;
    void *ctx;
    int32_t trace_on;
    uint64_t *trace;
    int32_t *rows;             /* NEWVOLT table row of each compartment in the last chip update, or NULL */
    /* synaptic events: the synchans' pending lists (ohs_set_synapses,
     * ohs_post_event) and the simulation time they are timed with */
    int32_t nsyn;              /* synchans with a pending list; 0: none */
    int32_t *syn_of_op;        /* ops index -> synchan id, -1 */
    ohs_pend *pend;            /* [nsyn] */
    int32_t *conn_off;         /* per ops index of a SPIKE_OP: first connection (CSR over ops indices) */
    int32_t *conn_dst, *conn_T;
    double *conn_wdt;
    double simtime;            /* SimulationTime() of the next step (sim_step.c:106-112) */
    double t_last;             /* SimulationTime() of the last step run */
};

static void *dupmem(const void *p, size_t n)
{
    void *q = calloc(n ? n : 1, 1);
    if (p && n) memcpy(q, p, n);
    return q;
}

ohs *ohs_create(const ohs_arrays *a)
{
    ohs *h = (ohs *)calloc(1, sizeof(ohs));
    h->a = *a;
    h->ops = (int32_t *)dupmem(a->ops, sizeof(int32_t) * a->nops);
    h->chip = (double *)dupmem(a->chip, sizeof(double) * a->nchips);
    h->vm = (double *)dupmem(a->vm, sizeof(double) * a->ncompts);
    /* two leading pad values: after row 0 the reference's backward pass loads
     * the entry before results[] (hines_solve.c:284-286), a value it never uses */
    h->results_base = (double *)calloc(2 * (size_t)a->ncompts + 4, sizeof(double));
    h->results = h->results_base + 2;
    h->ravals = (double *)dupmem(a->ravals, sizeof(double) * a->nravals);
    if (a->givals)
        h->givals = (double *)dupmem(a->givals, sizeof(double) * a->ngivals);
    else
        h->givals = NULL;
    if (a->nconcs > 0) {
        h->conc = (double *)dupmem(a->conc, sizeof(double) * (a->nconcs + 1));
        h->flux = (double *)calloc(a->nconcs + 1, sizeof(double));
        h->codiag = (double *)calloc(a->nconcs + 1, sizeof(double));
        h->coresult = (double *)calloc(a->nconcs + 1, sizeof(double));
        /* h_init_conc_solve sets every diagonal to 1 (hines_conc_init.c:127-131) */
        for (int32_t i = 0; i <= a->nconcs; i++) h->codiag[i] = 1.0;
    }
    if (a->ndiffs > 0) {
        size_t rows = 1, cells = 1;
        for (int32_t i = 0; i < a->ndiffs; i++) {
            const size_t r = (size_t)a->nshells[i] * (size_t)a->nshrows[i];
            if (r > rows) rows = r;
            if (r * (size_t)a->nshrows[i] > cells) cells = r * (size_t)a->nshrows[i];
        }
        h->dscr_cv = (double *)calloc(cells + 1, sizeof(double));
        h->dscr_rs = (double *)calloc(rows + 1, sizeof(double));
    }
    if (a->nelm_names > 0) h->itotal = (double *)calloc(a->nelm_names, sizeof(double));
    h->capspk = 1024;
    h->spike_op = (int32_t *)malloc(sizeof(int32_t) * h->capspk);
    h->spike_step = (int32_t *)malloc(sizeof(int32_t) * h->capspk);
    h->trace = (uint64_t *)calloc(a->ncompts + 1, sizeof(uint64_t));
    return h;
}

void ohs_destroy(ohs *h)
{
    if (!h) return;
    free(h->ops); free(h->chip); free(h->vm); free(h->results_base); free(h->ravals);
    free(h->givals); free(h->conc); free(h->flux); free(h->codiag); free(h->itotal);
    free(h->coresult); free(h->dscr_cv); free(h->dscr_rs);
    free(h->spike_op); free(h->spike_step); free(h->trace); free(h->rows);
    free(h->conn_off); free(h->conn_dst); free(h->conn_T); free(h->conn_wdt);
    if (h->pend)
        for (int32_t i = 0; i < h->nsyn; i++) free(h->pend[i].ev);
    free(h->pend); free(h->syn_of_op);
    free(h);
}

void ohs_set_hooks(ohs *h, ohs_syn_fn syn, ohs_rand_fn rnd, void *ctx)
{
    h->syn = syn; h->rnd = rnd; h->ctx = ctx;
}

void ohs_trace_enable(ohs *h, int32_t on) { h->trace_on = on; }

void ohs_rows_enable(ohs *h, int32_t on)
{
    free(h->rows);
    h->rows = NULL;
    if (on) {
        h->rows = (int32_t *)malloc(sizeof(int32_t) * (size_t)h->a.ncompts);
        for (int32_t i = 0; i < h->a.ncompts; i++) h->rows[i] = -1;
    }
}

void ohs_rows(ohs *h, int32_t *rows)
{
    if (h->rows) memcpy(rows, h->rows, sizeof(int32_t) * (size_t)h->a.ncompts);
}
void ohs_trace(ohs *h, uint64_t *per_comp)
{
    memcpy(per_comp, h->trace, sizeof(uint64_t) * h->a.ncompts);
}

static inline uint64_t fnv(uint64_t x, int32_t v)
{
    uint32_t u = (uint32_t)v;
    for (int i = 0; i < 4; i++) { x ^= (u >> (8 * i)) & 0xffu; x *= 1099511628211ull; }
    return x;
}

static void record_spike(ohs *h, int32_t opidx)
{
    if (h->nspk == h->capspk) {
        h->capspk *= 2;
        h->spike_op = (int32_t *)realloc(h->spike_op, sizeof(int32_t) * h->capspk);
        h->spike_step = (int32_t *)realloc(h->spike_step, sizeof(int32_t) * h->capspk);
    }
    h->spike_op[h->nspk] = opidx;
    h->spike_step[h->nspk] = h->step;
    h->nspk++;
}

/* EVENT of an hsolved synchan (newconn/synchan.c:161-302): insert an event
 * that hits `time` steps from now into the pending list at simulation time
 * t.  A non-empty list first has its head's time field re-timed by the whole
 * steps elapsed since time_last_event, (int)((t - tle)/dt) of the two
 * floating simulation times (:249-250); the new event goes before the first
 * event whose (relative) time is >= its own, becoming relative to the event
 * before it, and the event after it becomes relative to the new one
 * (:251-294).  A new head sets the SYN2_OP counter (h_eventaction ->
 * HSEVENT, hines.c:352-357).  time_last_event = t (:300). */
static void pend_insert(ohs *h, int32_t sid, int32_t time, double wdt, double t)
{
    ohs_pend *q = &h->pend[sid];
    if (q->n == q->cap) {
        q->cap = q->cap ? 2 * q->cap : 8;
        q->ev = (ohs_ev *)realloc(q->ev, sizeof(ohs_ev) * q->cap);
    }
    if (q->n == 0) {
        q->ev[0].time = time;
        q->ev[0].wdt = wdt;
        q->n = 1;
        h->ops[q->synop + 2] = time;
    } else {
        q->ev[0].time -= (int32_t)((t - q->tle) / h->a.dt);
        int32_t i = 0;
        for (; i < q->n; i++) {
            if (time <= q->ev[i].time) break;
            time -= q->ev[i].time;
        }
        if (i < q->n) {
            q->ev[i].time -= time;
            memmove(&q->ev[i + 1], &q->ev[i], sizeof(ohs_ev) * (size_t)(q->n - i));
        }
        q->ev[i].time = time;
        q->ev[i].wdt = wdt;
        q->n++;
        if (i == 0) h->ops[q->synop + 2] = time;
    }
    q->tle = t;
}

/* HPROCESS of a synchan whose SYN2_OP counter reached 0 (h_dosynchan,
 * hines_conn.c:60-89; newconn/synchan.c:305-358): the head's time is forced
 * to 0, the activation (a float field, struct_defs.h:112) restarts at 0 and
 * takes weight/dt of every event at the front whose time is 0, in list
 * order; a list left non-empty sets time_last_event = t.  Returns the
 * activation; the caller adds dt*activation*stablist[k+3] to X and sets the
 * counter to the new head's time or -1. */
static float pend_fire(ohs *h, int32_t sid, double t)
{
    ohs_pend *q = &h->pend[sid];
    float act = 0.0f;
    if (q->n == 0) return act;
    q->ev[0].time = 0;
    int32_t i = 0;
    while (i < q->n && q->ev[i].time == 0) {
        act = (float)((double)act + q->ev[i].wdt);
        i++;
    }
    memmove(&q->ev[0], &q->ev[i], sizeof(ohs_ev) * (size_t)(q->n - i));
    q->n -= i;
    if (q->n > 0) q->tle = t;
    return act;
}

/* A spike of the SPIKE_OP at ops index op: the spikegen's messages in
 * connection order, each an EVENT at its synchan (h_dospike_event,
 * hines_conn.c:44-56). */
static void deliver(ohs *h, int32_t op)
{
    for (int32_t j = h->conn_off[op]; j < h->conn_off[op + 1]; j++)
        pend_insert(h, h->conn_dst[j], h->conn_T[j], h->conn_wdt[j], h->simtime);
}

/* The pending list of the synchan whose SYN2_OP is at ops index op, made on
 * first use; -1 if ops[op] is not a SYN2_OP. */
static int32_t syn_id(ohs *h, int32_t op)
{
    const int32_t nops = h->a.nops;
    if (!h->syn_of_op) {
        h->syn_of_op = (int32_t *)malloc(sizeof(int32_t) * (size_t)nops);
        for (int32_t i = 0; i < nops; i++) h->syn_of_op[i] = -1;
    }
    if (op < 0 || op + 3 >= nops || h->ops[op] != SYN2_OP) return -1;
    if (h->syn_of_op[op] < 0) {
        h->pend = (ohs_pend *)realloc(h->pend, sizeof(ohs_pend) * (size_t)(h->nsyn + 1));
        memset(&h->pend[h->nsyn], 0, sizeof(ohs_pend));
        h->pend[h->nsyn].synop = op;
        h->syn_of_op[op] = h->nsyn++;
    }
    return h->syn_of_op[op];
}

int32_t ohs_set_synapses(ohs *h, int32_t n, const int32_t *src_op, const int32_t *dst_op,
                         const double *weight, const double *delay)
{
    const int32_t nops = h->a.nops;
    const double dt = h->a.dt;
    free(h->conn_off); free(h->conn_dst); free(h->conn_T); free(h->conn_wdt);
    h->conn_off = NULL; h->conn_dst = h->conn_T = NULL; h->conn_wdt = NULL;
    if (n <= 0) return 0;
    int32_t mind = 1 << 30;
    int32_t *T = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t i = 0; i < n; i++) {
        if (src_op[i] < 0 || src_op[i] >= nops || h->ops[src_op[i]] != SPIKE_OP || syn_id(h, dst_op[i]) < 0) {
            free(T);
            return -1;
        }
        /* delay is a float field of the synapse; the event hits
         * (int)((delay - (t - eventtime))/dt + 0.9) steps later, and an hsolve
         * spike is sent at the current time (newconn/synchan.c:193-194) */
        const double dl = (double)(float)delay[i];
        T[i] = (int32_t)((dl - 0.0) / dt + 0.9);
        if (T[i] < mind) mind = T[i];
    }
    h->conn_off = (int32_t *)calloc((size_t)nops + 1, sizeof(int32_t));
    h->conn_dst = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    h->conn_T = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    h->conn_wdt = (double *)malloc(sizeof(double) * (size_t)n);
    for (int32_t i = 0; i < n; i++) h->conn_off[src_op[i] + 1]++;
    for (int32_t i = 0; i < nops; i++) h->conn_off[i + 1] += h->conn_off[i];
    int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * ((size_t)nops + 1));
    memcpy(cur, h->conn_off, sizeof(int32_t) * ((size_t)nops + 1));
    for (int32_t i = 0; i < n; i++) {
        const int32_t q = cur[src_op[i]]++;
        h->conn_dst[q] = h->syn_of_op[dst_op[i]];
        h->conn_T[q] = T[i];
        /* synapse weight (float) / dt (newconn/synchan.c:351) */
        h->conn_wdt[q] = (double)(float)weight[i] / dt;
    }
    free(cur);
    free(T);
    return mind;
}

int32_t ohs_post_event(ohs *h, int32_t syn_op, int32_t time, double weight, int32_t after)
{
    const int32_t sid = syn_id(h, syn_op);
    if (sid < 0 || time < 0) return -1;
    if (after && h->step == 0) return -1;
    pend_insert(h, sid, time, (double)(float)weight / h->a.dt, after ? h->t_last : h->simtime);
    return 0;
}

/* h_in_msgs (hines_msgs.c:351-405).  The activation is scaled by the full
 * dt even under Crank-Nicolson (the local dt = dt/2 is not used there). */
int32_t ohs_in_msgs(ohs *h, int32_t n, const ohs_input *in)
{
    for (int32_t i = 0; i < n; i++) {
        const ohs_input *e = &in[i];
        switch (e->op) {
        case 1:   /* INJECT_OP */
        case 2:   /* BASE_OP */
        case 3:   /* THRESH_OP */
            if (e->index < 0 || e->index >= h->a.nchips) return -1;
            h->chip[e->index] = e->value;
            break;
        case SYN2_OP:
            if (e->index < 0 || e->index >= h->a.nchips || e->stab < 0 || e->stab + 3 >= 4 * h->a.nstab) return -1;
            if (e->value > 1.0e-10) {            /* VTINY, hines_defs.h:115 */
                if (e->per_dt)
                    h->chip[e->index] += e->value * h->a.stablist[e->stab + 3] * h->a.dt;
                else
                    h->chip[e->index] += e->value * h->a.stablist[e->stab + 3];
            }
            break;
        case FLUX_OP:
            if (e->index < 0 || e->index >= h->a.nconcs || !h->flux) return -1;
            h->flux[e->index] = e->value;
            break;
        default:
            return -1;
        }
    }
    return 0;
}

void ohs_set_time(ohs *h, double t) { h->simtime = t; }
double ohs_time(ohs *h) { return h->simtime; }

int32_t ohs_pending(ohs *h, int32_t syn_op, int32_t *time, double *wdt, int32_t cap, double *tle)
{
    if (!h->syn_of_op || syn_op < 0 || syn_op >= h->a.nops || h->syn_of_op[syn_op] < 0) return -1;
    const ohs_pend *q = &h->pend[h->syn_of_op[syn_op]];
    for (int32_t i = 0; i < q->n && i < cap; i++) {
        if (time) time[i] = q->ev[i].time;
        if (wdt) wdt[i] = q->ev[i].wdt;
    }
    if (tle) *tle = q->tle;
    return q->n;
}

/* do_chip_hh4_update with SOLVEINTERPOL (hines_chip.c:78-602).  The op
 * cursor `p`, the chip cursor `c` and the givals cursor `g` advance exactly as
 * the reference's pointers do. */
int32_t ohs_chip_update(ohs *h)
{
    const ohs_arrays *A = &h->a;
    int32_t *ops = h->ops;
    double *chip = h->chip;
    double *results = h->results;
    const double *vmv = h->vm;
    const double *tablist = A->tablist, *xvals = A->xvals, *yvals = A->yvals;
    const double *stablist = A->stablist;
    double *conc = h->conc, *flux = h->flux, *codiag = h->codiag, *itotal = h->itotal;
    const int ncols = A->ncols, xdivs = A->xdivs, ydivs = A->ydivs;
    const double xmin = A->xmin, invdx = A->invdx, ymin = A->ymin, invdy = A->invdy;
    const int chanmode5 = (A->chanmode == 5);
    const int storeall = A->storemode;
    const int keepg = (h->givals != NULL);
    double *givals = keepg ? h->givals + 1 : NULL;
    double gscratch[16];

    size_t p = 0, c = 0, r = 0, vi = 0;
    int k = 0, cfilo = 0, filo = 0, doGHK = 0;
    double Vm, sumgchan, ichan, Gk = 0, X, Y, Ek = 0.0, GHK = 0.0, concen, Ik = 0.0;
    double vipol = 0.0, cipol = 0.0, ipol = 0.0, Im = 0.0, gmax = 0.0;
    double xlo, xhi, ylo = 0.0, yhi = 0.0;
    const double *vtab = NULL, *ctab = NULL, *tab = NULL, *tabn;
    double *concin = NULL, *fluxindex = NULL, *concout;
    uint64_t tr = 0;
    int comp = -1;

    if (tablist) { xlo = xhi = 0.0; } else { xlo = -1e20; xhi = 1e20; }
    if (storeall && itotal)
        for (k = 0; k < A->nelm_names; k++) itotal[k] = 0.0;

#define GOUT(v) do { if (keepg) *givals++ = (v); } while (0)

    while (ops[p++] != LCOMPT_OP) {
        /* start of a compartment (hines_chip.c:157-168) */
        if (h->trace_on && comp >= 0) h->trace[comp] = tr;
        comp++;
        tr = 1469598103934665603ull;
        Vm = vmv[vi++];
        sumgchan = 0.0;
        ichan = chip[c] + chip[c + 1];
        if (keepg) {
            Im = chip[c] - Vm / *givals++;
            *givals++ = Im;
        } else {
            Im = 0.0;
        }
        if (storeall == 1 && itotal) itotal[0] += Im;
        c += 2;
        (void)gscratch;

        while (ops[p] > LCOMPT_OP) {
            int op = ops[p++];
            if (h->trace_on) tr = fnv(tr, op);
            switch (op) {
            case NEWVOLT_OP:
                vipol = (Vm - xmin) * invdx;
                if ((xlo > Vm) || (Vm >= xhi)) {
                    if (vipol < 0.0) vipol = 0.0;
                    else if (vipol > xdivs) vipol = xdivs;
                    filo = (int)vipol;
                    vipol = vipol - (double)filo;
                    xlo = xvals[filo];
                    xhi = xvals[filo + 1];
                    vtab = tablist + (size_t)filo * ncols;
                } else {
                    vipol = vipol - (double)filo;
                }
                if (h->rows) h->rows[comp] = filo;
                continue;
            case CONC_VAL_OP:
                k = ops[p++];
                if (h->trace_on) tr = fnv(tr, k);
                concin = conc + k;
                fluxindex = flux + k;
                *fluxindex = 0.0;
                continue;
            case DOMAIN_OP:
                concen = chip[c++] * Ik;
                goto cinterpol;
            case NEWCONC2_OP:
                if (h->trace_on) tr = fnv(tr, ops[p]);
                concen = conc[ops[p++]];
                tab = ctab;
                ipol = cipol;
                goto cinterpol;
            case NEWCONC1_OP:
                concen = *concin;
            cinterpol:
                cipol = (concen - ymin) * invdy;
                if ((ylo > concen) || (concen >= yhi)) {
                    if (cipol < 0.0) cipol = 0.0;
                    else if (cipol > ydivs) cipol = ydivs;
                    cfilo = (int)cipol;
                    cipol = cipol - (double)cfilo;
                    ylo = yvals[cfilo];
                    yhi = yvals[cfilo + 1];
                    ctab = tablist + (size_t)cfilo * ncols;
                } else {
                    cipol = cipol - cfilo;
                }
                continue;
            case SPIKE_OP:
                ops[p] -= 1;
                if (Vm > chip[c++]) {
                    if (ops[p] <= 0) {
                        record_spike(h, (int32_t)p - 1);
                        if (h->conn_off) deliver(h, (int32_t)p - 1);
                        ops[p] = ops[p + 1];
                    }
                }
                p += 2;
                continue;
            case ELMMPUMP_OP:
                X = 1.0 / (*concin + chip[c++]);
                ichan += Ik = chip[c] * *concin * X;
                if (chanmode5) GOUT(Ik / chip[c]);
                else GOUT(Ik);
                c++;
                Im += Ik;
                codiag[k] += chip[c++] * X;
                continue;
            case FLUX_OP:
                *fluxindex += Ik;
                continue;
            case STORE_ALLI_OP:
                if (h->trace_on) tr = fnv(tr, ops[p]);
                itotal[ops[p++]] += Ik;
                continue;
            case STORE_ALLC_OP:
                if (h->trace_on) tr = fnv(tr, ops[p]);
                itotal[ops[p++]] += Gk;
                continue;
            case SCALE_FLUX_OP:
                *fluxindex += chip[c++] * Ik;
                continue;
            case GHK_OP:
                concout = concin + 1;
                GHK = chip[c++];
                X = Vm * GHK;
                Y = exp(-X);
                Ek = chip[c++];
                Ik = Ek * (*concin - (Y * *concout));
                if (fabs(X) < OHS_TINY) {
                    Gk = 1.0 - 0.5 * X;
                    GHK = 0.5 * GHK * Ek * (*concin + *concout);
                    Ik = -Ik / Gk;
                } else {
                    Gk = 1.0 - Y;
                    GHK = GHK * (Ik / Gk - Ek * GHK * (*concin - *concout) * Vm * Y / (Gk * Gk));
                    Ik = -X * Ik / Gk;
                }
                Ek = (Ik + GHK * Vm) / GHK;
                /* fall through */
            case SAME_GHK_OP:
                gmax = chip[c++];
                Gk = gmax * GHK;
                break;
            case DO_TABGHK_OP:
                doGHK = 1;
                break;
            case NERNST_OP:
                Ek = log(*(concin + 1) / *concin) * chip[c++];
                gmax = Gk = chip[c++];
                break;
            case CHAN_EK_OP:
                Ek = chip[c++];
                /* fall through */
            case CHAN_OP:
                gmax = Gk = chip[c++];
                break;
            default:
                return -1;
            }

            /* conductance ops until the current is added (hines_chip.c:404-590) */
            for (;;) {
                k = ops[p++];
                if (h->trace_on) tr = fnv(tr, k);
                if (k == IPOL1V_OP || k == IPOL1C_OP) {
                    double f = (k == IPOL1V_OP) ? vipol : cipol;
                    if (h->trace_on) tr = fnv(tr, ops[p]);
                    tab = ((k == IPOL1V_OP) ? vtab : ctab) + ops[p++];
                    X = *tab;
                    tabn = tab + ncols;
                    X = X + f * (*tabn - X);
                    tab++;
                    tabn++;
                    Y = *tab;
                    Y = Y + f * (*tabn - Y);
                    goto dopower;
                } else if (k == IPOL2V_OP || k == IPOL2C_OP) {
                    double f;
                    if (h->trace_on) tr = fnv(tr, ops[p]);
                    if (k == IPOL2V_OP) {
                        tab = vtab + cfilo + cfilo + ops[p++];
                        f = vipol;
                    } else {
                        tab += cfilo + cfilo + ops[p++];
                        f = ipol;
                    }
                    tabn = tab + ncols;
                    X = (1.0 - f) * ((1.0 - cipol) * (*tab) + cipol * (*(tab + 2))) +
                        f * ((1.0 - cipol) * (*tabn) + cipol * (*(tabn + 2)));
                    tab++;
                    tabn++;
                    Y = (1.0 - f) * ((1.0 - cipol) * (*tab) + cipol * (*(tab + 2))) +
                        f * ((1.0 - cipol) * (*tabn) + cipol * (*(tabn + 2)));
                dopower:
                    k = ops[p++];
                    if (h->trace_on) tr = fnv(tr, k);
                    if (k == 0) {
                        Ek = Y;
                        if (doGHK) {
                            GHK = X;
                            doGHK = 0;
                            break;
                        } else {
                            Gk *= X;
                            goto doaddcurr;
                        }
                    }
                    if (k > 0) {
                        X = chip[c] = (chip[c] * (2.0 - X) + Y) / X;
                    } else {
                        X = chip[c] = Y / X;
                        k = -k;
                    }
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
                } else if (k >= SYN2_OP && k <= SYN3_OP) {
                    const size_t synop = p - 1;
                    if (k == SYN3_OP) {
                        double u = h->rnd ? h->rnd(h->ctx) : 1.0;
                        if (u <= chip[c++]) chip[c] += stablist[ops[p] + 3];
                    }
                    k = ops[p++];
                    if (h->trace_on) tr = fnv(tr, k);
                    chip[c] *= stablist[k];
                    if (ops[p] == 0) {
                        int32_t nxt = -1;
                        const int32_t sid = h->syn_of_op ? h->syn_of_op[synop] : -1;
                        if (sid >= 0) {
                            /* h_dosynchan (hines_conn.c:60-89) */
                            const float act = pend_fire(h, sid, h->simtime);
                            chip[c] += A->dt * (double)act * stablist[k + 3];
                            nxt = h->pend[sid].n > 0 ? h->pend[sid].ev[0].time : -1;
                        } else if (h->syn) {
                            nxt = h->syn(h->ctx, k, ops[p + 1], (int32_t)c, h->step);
                        }
                        ops[p] = nxt;
                    }
                    ops[p] -= 1;
                    p += 2;
                    X = chip[c];
                    c++;
                    Gk *= (chip[c] = X * stablist[k + 1] + chip[c] * stablist[k + 2]);
                    c++;
                    if (ops[p] == MGBLOCK_OP) {
                        if (chanmode5) {
                            GOUT(Gk / gmax);
                            GOUT(Gk * (Ek - Vm) / gmax);
                        } else {
                            GOUT(Gk);
                            GOUT(Gk * (Ek - Vm));
                        }
                        GOUT(Ek);
                        X = chip[c] * exp(Vm * chip[c + 1]);
                        c += 2;
                        Gk *= X / (X + chip[c++]);
                        p++;
                    }
                    goto doaddcurr;
                } else if (k == ADD_CURR_OP) {
                doaddcurr:
                    sumgchan += Gk;
                    ichan += Ek * Gk;
                    if (chanmode5) {
                        GOUT(Gk / gmax);
                        Ik = Gk * (Ek - Vm);
                        GOUT(Ik / gmax);
                    } else {
                        GOUT(Gk);
                        Ik = Gk * (Ek - Vm);
                        GOUT(Ik);
                    }
                    GOUT(Ek);
                    Im += Ik;
                    break;
                } else {
                    return -1;
                }
            }
        }
        /* end of compartment (hines_chip.c:592-599) */
        results[r++] = Vm + ichan * chip[c];
        results[r++] = sumgchan * chip[c] + chip[c + 1];
        c += 2;
        GOUT(Im);
    }
    if (h->trace_on && comp >= 0) h->trace[comp] = tr;
#undef GOUT
    return 0;
}

/* Shared forward sweep of do_euler_hsolve / do_crank_hsolve
 * (hines_solve.c:236-273 and :153-190).  Returns the funcs cursor after the
 * forward FINISH and leaves the row cursor on the last row. */
static size_t fwd_sweep(ohs *h, size_t *rowp, double *rv, double *dv, size_t *rap)
{
    const int32_t *funcs = h->a.funcs;
    double *res = h->results;
    double *raval = h->ravals;
    size_t f = 0, ra = 0, row = 2;
    double temp = 0.0, resultval = res[row], diaval = res[row + 1];
    int op = funcs[f++];
    for (;;) {
        if (op == FORWARD_ELIM) {
            temp = raval[ra++] / res[funcs[f] + 1];
            diaval -= raval[ra++] * temp;
            resultval -= res[funcs[f++]] * temp;
        } else if (op == SET_DIAG) {
            if (row + 3 >= 2 * (size_t)h->a.ncompts) { h->bad_funcs = 1; break; }
            res[row] = resultval;
            res[row + 1] = diaval;
            row += 2;
            resultval = res[row];
            diaval = res[row + 1];
        } else if (op == SKIP_DIAG) {
            if (row + 5 >= 2 * (size_t)h->a.ncompts) { h->bad_funcs = 1; break; }
            res[row] = resultval;
            res[row + 1] = diaval;
            row += 4;
            resultval = res[row];
            diaval = res[row + 1];
        } else if (op == FASTSIBARRAY_ELIM) {
            raval[funcs[f++]] -= raval[ra++] * temp;
        } else if (op == COPY_ARRAY) {
            raval[funcs[f++]] = raval[ra++];
        } else if (op == SIBARRAY_ELIM) {
            raval[funcs[f + 1]] -= raval[funcs[f]] * temp;
            f += 2;
        } else {
            break;
        }
        op = funcs[f++];
    }
    *rowp = row;
    *rv = resultval;
    *dv = diaval;
    *rap = ra;
    return f;
}

void ohs_euler_solve(ohs *h)
{
    double *res = h->results;
    double *vm = h->vm;
    const int32_t *funcs = h->a.funcs;
    const double *raval = h->ravals;
    int n = h->a.ncompts;
    if (n == 1) {
        vm[0] = res[0] / res[1];
        return;
    }
    size_t row, ra;
    double resultval, diaval;
    size_t f = fwd_sweep(h, &row, &resultval, &diaval, &ra);
    size_t vi = (size_t)n - 1;
    if (h->bad_funcs || row != 2 * vi) { h->bad_funcs = 1; return; }
    vm[vi] = res[row] = resultval / diaval;
    row -= 2;
    resultval = res[row];
    for (;;) {
        int op = funcs[f++];
        if (op == BACKWARD_ELIM) {
            resultval -= raval[ra++] * res[funcs[f++]];
        } else if (op == CALC_RESULTS) {
            if (vi == 0) { h->bad_funcs = 1; break; }
            vi--;
            vm[vi] = res[row] = resultval / res[row + 1];
            row -= 2;
            resultval = res[row];
        } else if (op == SIBARRAY_ELIM) {
            resultval -= raval[funcs[f]] * res[funcs[f + 1]];
            f += 2;
        } else {
            break;
        }
    }
}

void ohs_crank_solve(ohs *h)
{
    double *res = h->results;
    double *vm = h->vm;
    const int32_t *funcs = h->a.funcs;
    const double *raval = h->ravals;
    int n = h->a.ncompts;
    if (n == 1) {
        double x = res[0] / res[1];
        vm[0] = x + x - vm[0];
        return;
    }
    size_t row, ra;
    double resultval, diaval;
    size_t f = fwd_sweep(h, &row, &resultval, &diaval, &ra);
    size_t vi = (size_t)n - 1;
    if (h->bad_funcs || row != 2 * vi) { h->bad_funcs = 1; return; }
    resultval = resultval / diaval;
    vm[vi] = resultval + resultval - vm[vi];
    res[row] = resultval;
    row -= 2;
    resultval = res[row];
    for (;;) {
        int op = funcs[f++];
        if (op == BACKWARD_ELIM) {
            resultval -= raval[ra++] * res[funcs[f++]];
        } else if (op == CALC_RESULTS) {
            if (vi == 0) { h->bad_funcs = 1; break; }
            resultval = resultval / res[row + 1];
            vi--;
            vm[vi] = resultval + resultval - vm[vi];
            res[row] = resultval;
            row -= 2;
            resultval = res[row];
        } else if (op == SIBARRAY_ELIM) {
            resultval -= raval[funcs[f]] * res[funcs[f + 1]];
            f += 2;
        } else {
            break;
        }
    }
}

/* X^(hill-1) as the Hill pumps form it (hines_conc_chip.c:126-135) */
static double hill_pow(int32_t hill, double X)
{
    if (hill == 2) return X;
    if (hill == 3) return X * X;
    if (hill == 4) return X * X * X;
    return 1.0;
}

/* The concentration program (do_h4_conc_chip_update, hines_conc_chip.c:59-222,
 * built with CHANMODE4: FLUXCONC_OP and CACONC_OP leave the flux alone, the
 * chip pass's CONC_VAL_OP resets it).  Two registers run through it: the
 * concentration the ops apply to and the sum of their terms, which CONC_OP
 * (before moving on) and LCONC_OP store as that concentration's coresult.
 * The reference runs this pass after the whole chip update and before the
 * solves (hines.c:238-240).  TAUPUMP_OP's chanmode-4 copy of its rate into
 * the solver's ovals[] is not kept. */
int32_t ohs_conc_chip_update(ohs *h)
{
    const int32_t *op = h->a.concops;
    const double *cc = h->a.concchip;
    double *conc = h->conc, *flux = h->flux, *codiag = h->codiag, *cores = h->coresult;
    if (!op || h->a.nconcops <= 0) return 0;
    int32_t at = 0;
    double sum = 0.0;
    for (int32_t q = 0; q < h->a.nconcops;) {
        const int32_t code = op[q++];
        switch (code) {
        case CONC_OP:
            cores[at] = sum;
            /* fall through */
        case FCONC_OP:
            at = op[q++];
            sum = 0.0;
            break;
        case FLUXCONC_OP:               /* dt leak, dt/(z F vol) */
            sum += cc[0] + flux[at] * cc[1];
            cc += 2;
            break;
        case CTAUPUMP_OP:               /* Ceq dt/T_C, dt/T_C */
            sum += cc[0];
            codiag[at] += cc[1];
            cc += 2;
            break;
        case TAUPUMP_OP: {              /* T_C, T_A, T_V, 1/T_B, Ceq; operand: compartment */
            const double itau = h->a.dt / (cc[0] + cc[1] * exp((h->vm[op[q++]] - cc[2]) * cc[3]));
            sum += cc[4] * itau;
            codiag[at] += itau;
            cc += 5;
            break;
        }
        case MMPUMP_OP:                 /* dt vmax/vol, Kd */
            codiag[at] += cc[0] / (conc[at] + cc[1]);
            cc += 2;
            break;
        case HILLPUMP_OP: {             /* Kd, dt vmax/vol; operand: Hill coefficient */
            const double X = conc[at];
            double c = hill_pow(op[q++], X);
            c = c / (c * X + cc[0]);
            codiag[at] += cc[1] * c;
            cc += 2;
            break;
        }
        case PHILLPUMP_OP: {            /* Kd, dt vmax/vol shell, dt vmax/vol pool; operands: Hill, shell */
            const int32_t hill = op[q++];
            const int32_t sh = op[q++];
            const double X = conc[sh];
            double c = hill_pow(hill, X);
            c = c / (c * X + cc[0]);
            codiag[sh] += cc[1] * c;
            sum += cc[2] * c * X;
            cc += 3;
            break;
        }
        case BUFFER_OP: {               /* kb dt/2, dt kb Btot, dt kf; operand: the buffer */
            const int32_t b = op[q++];
            const double X = cc[0];
            const double c = ((1.0 - X) * conc[b] + cc[1]) / (1.0 + X + cc[2] * conc[at]);
            sum += c - conc[b];
            conc[b] = c;
            cc += 3;
            break;
        }
        case POOL_OP:
            conc[at] += sum;
            break;
        case CACONC_OP: {               /* base, 1 + dt/(2 tau), dt B; operand: the pool */
            const int32_t k = op[q++];
            const double base = cc[0];
            const double itau = cc[1];
            conc[k] = base + ((conc[k] - base) * (2.0 - itau) + (flux[k] * cc[2])) / itau;
            cc += 3;
            break;
        }
        case LCONC_OP:
            cores[at] = sum;
            return 0;
        case LCACONC_OP:
            return 0;
        default:
            return -1;
        }
    }
    return -1;
}

/* Shells alone (nshrows 1): a tridiagonal system per step, eliminated from
 * the innermost shell out and substituted back (hines_conc.c:322-393).  Row j
 * reads  (codiag_j + Ddn_j + Dup_j) x_j - Ddn_j x_(j-1) - Dup_j x_(j+1)
 *        = C_j + coresult_j
 * with the coupling values of covals0 in order (shell 0: up; inner shells:
 * down, up; the last: down).  Crank-Nicolson takes C <- 2x - C.  Every
 * diagonal used is set back to 1.  Returns the covals0 cursor after the
 * system. */
static int32_t conc_tridiag(ohs *h, int32_t c0, int32_t ms, int32_t l)
{
    double *conc = h->conc, *codiag = h->codiag, *cores = h->coresult;
    double *cv = h->dscr_cv, *rs = h->dscr_rs;
    const double *cov0 = h->a.covals0;
    double up = cov0[l++];
    double dg = up + codiag[c0];
    codiag[c0] = 1.0;
    rs[0] = (conc[c0] + cores[c0]) / dg;
    int32_t j;
    for (j = 1; j < ms - 1; j++) {
        const double e = cv[j] = -up / dg;
        const double dn = cov0[l];
        up = cov0[l + 1];
        l += 2;
        dg = codiag[c0 + j] + dn * (1.0 + e) + up;
        codiag[c0 + j] = 1.0;
        rs[j] = (conc[c0 + j] + cores[c0 + j] + dn * rs[j - 1]) / dg;
    }
    const double e = cv[j] = -up / dg;
    const double dn = cov0[l++];
    dg = codiag[c0 + j] + dn * (1.0 + e);
    codiag[c0 + j] = 1.0;
    double x = (conc[c0 + j] + cores[c0 + j] + dn * rs[j - 1]) / dg;
    if (h->a.method == CRANK_INT) {
        conc[c0 + j] = x + x - conc[c0 + j];
        for (int32_t r = j - 1; r >= 0; r--) {
            x = rs[r] - cv[r + 1] * x;
            conc[c0 + r] = x + x - conc[c0 + r];
        }
    } else {
        conc[c0 + j] = x;
        for (int32_t r = j - 1; r >= 0; r--) {
            x = rs[r] - cv[r + 1] * x;
            conc[c0 + r] = x;
        }
    }
    return l;
}

/* Shells with buffers (nshrows dw > 1): the banded Crank-Nicolson system of
 * hines_conc.c:106-321, filled and forward-eliminated shell by shell from
 * the innermost out, then substituted back.  Rows are numbered r = m*dw + n
 * (shell m, row n: the nd fixed buffers, the diffusible buffers, the shell
 * last); the scratch holds per row r the dw eliminated upper entries at
 * cv[r*dw ...] and the eliminated right-hand side rs[r].  Buffer rate
 * triples (dt/2 kf, dt/2 kb, dt kb Btot) come from bufvals, per-shell totals
 * of fixed buffers and every coupling from covals0 in the order the rows use
 * them.  The reference computes a diffusible buffer row's upper entry twice
 * (hines_conc.c:228-232, then :250); only the second value is kept, so only
 * it is computed here.  Returns the covals0 cursor after the system. */
static int32_t conc_banded(ohs *h, int32_t c0, int32_t ms, int32_t dw, int32_t nd, int32_t nb0,
                           int32_t l)
{
    double *conc = h->conc, *codiag = h->codiag, *cores = h->coresult;
    double *cv = h->dscr_cv, *rs = h->dscr_rs;
    const double *cov0 = h->a.covals0, *bv = h->a.bufvals;
    const int32_t sw = dw * dw;
    double co = conc[c0 + dw - 1];            /* the shell's concentration at t */
    for (int32_t m = 0; m < ms; m++) {
        const int32_t rS = m * dw + dw - 1;   /* the shell row */
        const int32_t xS = rS * dw;           /* its scratch entries */
        const int32_t xP = xS - sw;           /* the previous shell row's */
        const int32_t rP = rS - dw;
        const int last = (m == ms - 1);
        double dgc = codiag[c0 + rS];
        codiag[c0 + rS] = 1.0;
        double rsc = cores[c0 + rS];
        for (int32_t n = 0; n < nd; n++) {    /* fixed buffers */
            const int32_t r = m * dw + n, nb = nb0 + 3 * n;
            double dg = cv[xS + n] = bv[nb] * co + bv[nb + 1];
            const double t = bv[nb] * conc[c0 + r];
            dgc += t;
            dg = 1.0 / (1.0 + dg);
            cv[r * dw] = t * dg;
            const double t2 = -bv[nb + 1] * conc[c0 + r] + cov0[l++];
            rsc += t2;
            rs[r] = (conc[c0 + r] + t2) * dg;
        }
        for (int32_t n = nd; n < dw - 1; n++) {   /* diffusible buffers */
            const int32_t r = m * dw + n, nb = nb0 + 3 * n;
            double dg = cv[xS + n] = bv[nb] * co + bv[nb + 1];
            const double t = bv[nb] * conc[c0 + r];
            dgc += t;
            double t2 = -bv[nb + 1] * conc[c0 + r] + bv[nb + 2];
            rsc += t2;
            if (m) {
                const double nw = cov0[l] * cv[(r - dw) * dw];
                if (!last) {
                    dg += 1.0 + cov0[l + 1] + cov0[l] * (1.0 + cv[(r - dw) * dw + 1]) - nw * cv[xP + n];
                    cv[r * dw + 1] = -cov0[l + 1] / dg;
                    t2 += cov0[l] * conc[c0 + r - dw] + (1.0 - cov0[l] - cov0[l + 1]) * conc[c0 + r] +
                          cov0[l + 1] * conc[c0 + r + dw];
                } else {
                    dg += 1.0 + cov0[l] * (1.0 + cv[(r - dw) * dw + 1]) - nw * cv[xP + n];
                    t2 += cov0[l] * conc[c0 + r - dw] + (1.0 - cov0[l]) * conc[c0 + r];
                }
                rs[r] = (t2 + cov0[l] * rs[r - dw] - nw * rs[rP]) / dg;
                l += last ? 1 : 2;
            } else {
                dg += 1.0 + cov0[l];
                cv[r * dw + 1] = -cov0[l] / dg;
                rs[r] = (t2 + (1.0 - cov0[l]) * conc[c0 + r] + cov0[l] * conc[c0 + r + dw]) / dg;
                l++;
            }
            cv[r * dw] = t / dg;
        }
        /* the shell row: diffusion to the neighbours */
        if (m) {
            if (!last) {
                dgc += cov0[l + 1] + cov0[l] * (1.0 + cv[xP + dw - 1]);
                rsc += cov0[l] * (conc[c0 + rS - dw] + rs[rP]) + (1.0 - cov0[l] - cov0[l + 1]) * co;
                co = conc[c0 + rS + dw];
                rsc += cov0[l + 1] * co;
            } else {
                dgc += cov0[l] * (1.0 + cv[xP + dw - 1]);
                rsc += cov0[l] * (conc[c0 + rS - dw] + rs[rP]) + (1.0 - cov0[l]) * co;
            }
        } else {
            dgc += cov0[l];
            rsc += (1.0 - cov0[l]) * co;
            co = conc[c0 + rS + dw];
            rsc += cov0[l] * co;
        }
        /* eliminate the buffer entries of the shell row */
        for (int32_t n = 0; n < nd; n++) {
            const int32_t rb = m * dw + n;
            dgc -= cv[xS + n] * cv[rb * dw];
            rsc -= cv[xS + n] * rs[rb];
        }
        for (int32_t n = nd; n < dw - 1; n++) {
            const int32_t rb = m * dw + n;
            const double nw = m ? cv[xS + n] + cov0[l] * cv[xP + n] : cv[xS + n];
            dgc -= nw * cv[rb * dw];
            rsc -= nw * rs[rb];
            if (!last) cv[xS + n] *= -cv[rb * dw + 1];
        }
        if (!last) {
            for (int32_t n = nd; n < dw - 1; n++) cv[xS + n] /= dgc;
            if (m) {
                cv[xS + dw - 1] = -cov0[l + 1] / dgc;
                l += 2;
            } else {
                cv[xS + dw - 1] = -cov0[l] / dgc;
                l++;
            }
            rs[rS] = rsc / dgc;
        } else {
            co = conc[c0 + rS] = rsc / dgc;
            l++;
        }
    }
    /* back substitution, outermost shell first */
    for (int32_t m = ms - 1; m >= 0; m--) {
        const int32_t rS = m * dw + dw - 1;
        const int32_t xS = rS * dw;
        if (m < ms - 1) {
            co = rs[rS] - cv[xS + dw - 1] * conc[c0 + rS + dw];
            for (int32_t n = nd; n < dw - 1; n++) co -= cv[xS + n] * conc[c0 + rS + n + 1];
            conc[c0 + rS] = co;
            for (int32_t n = dw - 2; n >= nd; n--) {
                const int32_t rb = m * dw + n;
                conc[c0 + rb] = rs[rb] - cv[rb * dw] * co - cv[rb * dw + 1] * conc[c0 + rb + dw];
            }
        } else {
            for (int32_t n = dw - 2; n >= nd; n--) {
                const int32_t rb = m * dw + n;
                conc[c0 + rb] = rs[rb] - cv[rb * dw] * co;
            }
        }
        for (int32_t n = nd - 1; n >= 0; n--) {
            const int32_t rb = m * dw + n;
            conc[c0 + rb] = rs[rb] - cv[rb * dw] * co;
        }
    }
    return l;
}

/* do_h_conc_solve (hines_conc.c:39-396): every diffusion system in order,
 * one covals0 cursor running through them all.  Systems of fewer than two
 * shells are not stepped: the reference's elimination of one shell reads the
 * row after the system (its last-row step runs unconditionally). */
int32_t ohs_conc_solve(ohs *h)
{
    int32_t l = 0;
    for (int32_t i = 0; i < h->a.ndiffs; i++) {
        const int32_t ms = h->a.nshells[i], dw = h->a.nshrows[i];
        if (ms < 2 || dw < 1) return -1;
        if (dw == 1)
            l = conc_tridiag(h, h->a.dconc[i], ms, l);
        else
            l = conc_banded(h, h->a.dconc[i], ms, dw, h->a.nshfix[i], h->a.fbufval[i], l);
    }
    return 0;
}

void ohs_conc_regs(ohs *h, double *codiag, double *coresult)
{
    if (h->a.nconcs <= 0) return;
    if (codiag) memcpy(codiag, h->codiag, sizeof(double) * h->a.nconcs);
    if (coresult) memcpy(coresult, h->coresult, sizeof(double) * h->a.nconcs);
}

void ohs_get_conc(ohs *h, double *conc)
{
    if (h->a.nconcs > 0) memcpy(conc, h->conc, sizeof(double) * h->a.nconcs);
}

void ohs_put_conc(ohs *h, const double *conc)
{
    if (h->a.nconcs > 0) memcpy(h->conc, conc, sizeof(double) * h->a.nconcs);
}

int32_t ohs_step(ohs *h, int32_t nsteps)
{
    for (int32_t s = 0; s < nsteps; s++) {
        if (ohs_chip_update(h)) return -1;
        if (ohs_conc_chip_update(h)) return -1;
        if (h->a.ndiffs > 0 && ohs_conc_solve(h)) return -1;
        if (h->a.method == CRANK_INT) ohs_crank_solve(h);
        else ohs_euler_solve(h);
        if (h->bad_funcs) return -1;           /* malformed funcs[]: stopped inside the rows */
        h->step++;
        /* simulation_time += ClockValue(0) after every step (sim_step.c:112) */
        h->t_last = h->simtime;
        h->simtime += h->a.dt;
    }
    return 0;
}

int32_t ohs_stepno(ohs *h) { return h->step; }

void ohs_get(ohs *h, double *vm, double *chip, double *results, int32_t *ops, double *givals)
{
    if (vm) memcpy(vm, h->vm, sizeof(double) * h->a.ncompts);
    if (chip) memcpy(chip, h->chip, sizeof(double) * h->a.nchips);
    if (results) memcpy(results, h->results, sizeof(double) * 2 * h->a.ncompts);
    if (ops) memcpy(ops, h->ops, sizeof(int32_t) * h->a.nops);
    if (givals && h->givals) memcpy(givals, h->givals, sizeof(double) * h->a.ngivals);
}

void ohs_put(ohs *h, const double *vm, const double *chip, const int32_t *ops)
{
    if (vm) memcpy(h->vm, vm, sizeof(double) * h->a.ncompts);
    if (chip) memcpy(h->chip, chip, sizeof(double) * h->a.nchips);
    if (ops) memcpy(h->ops, ops, sizeof(int32_t) * h->a.nops);
}

int32_t ohs_nspikes(ohs *h) { return h->nspk; }
void ohs_spikes(ohs *h, int32_t *opidx, int32_t *step)
{
    memcpy(opidx, h->spike_op, sizeof(int32_t) * h->nspk);
    memcpy(step, h->spike_step, sizeof(int32_t) * h->nspk);
}
void ohs_clear_spikes(ohs *h) { h->nspk = 0; }

typedef struct {
    ohs **hs;
    int32_t nh, nsteps, nthreads, tid, rc;
} par_job;

static void *par_run(void *arg)
{
    par_job *j = (par_job *)arg;
    for (int32_t i = j->tid; i < j->nh; i += j->nthreads)
        if (ohs_step(j->hs[i], j->nsteps)) j->rc = -1;
    return NULL;
}

int32_t ohs_step_parallel(ohs **hs, int32_t nh, int32_t nsteps, int32_t nthreads)
{
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    par_job *jobs = (par_job *)calloc(nthreads, sizeof(par_job));
    for (int32_t t = 0; t < nthreads; t++) {
        jobs[t].hs = hs; jobs[t].nh = nh; jobs[t].nsteps = nsteps;
        jobs[t].nthreads = nthreads; jobs[t].tid = t;
        pthread_create(&th[t], NULL, par_run, &jobs[t]);
    }
    int32_t rc = 0;
    for (int32_t t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        if (jobs[t].rc) rc = -1;
    }
    free(th);
    free(jobs);
    return rc;
}
