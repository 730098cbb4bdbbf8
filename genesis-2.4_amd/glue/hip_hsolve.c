/*
 * hip_hsolve.c -- GENESIS glue for libhsolve_hip (compiled inside a GENESIS
 * tree, genesis/src/hines/hip/, with -DUSE_HIP; it includes GENESIS headers
 * and is not part of libhsolve_hip).
 *
 * Control flow follows the CUDA glue's (genesis/src/hines/cuda/
 * cuda_hsolve.c:289-517): lazy creation on the first PROCESS call, refusal
 * -> CPU path for this solver with the opcode list on stderr, one device
 * step per PROCESS.  The device keeps the state between steps.  What the
 * reference's PROCESS feeds a solver from outside reaches the device:
 *
 *   h_in_msgs (hines_msgs.c:351-405)  -> hsh_set_inputs_static each step,
 *                                        values read from the ininfo msgins;
 *   spikegen -> synchan of the solver -> a device connection table
 *                                        (hsh_set_synapses), the pending
 *                                        lists kept on the device;
 *   EVENTs from anywhere else        -> hip_synchan_event, called from the
 *                                        synchan EVENT action
 *                                        (newconn/synchan.c:161-302) for a
 *                                        device-run synchan, posted with
 *                                        hsh_post_events;
 *   spikes to elements outside       -> CallEventAction on those messages
 *                                        only, after the step (the device
 *                                        delivers the rest);
 *   HPUT, HCHANGE, HRESTORE, RESTORE2 -> hip_push_state after the host edit;
 *   HGET, HSAVE, SAVE2, out-msgs     -> hip_sync_chip before the host read;
 *   RESET                            -> hip_release (the next PROCESS
 *                                        creates the device state afresh).
 *
 * Solvers of one cell each (the reference's VAnet2 layout, thousands of
 * them) are merged into one group handle once the first step has shown
 * which solvers run and in which order: from the second step on, the first
 * member's PROCESS advances all members in one launch (hsh_create_group),
 * spikes between members are delivered on the device, and the other
 * members' PROCESS calls only pull their outputs.  As in the reference's
 * CUDA batching (cuda_hsolve.c:91-122), a member's message inputs are read
 * when the group steps, i.e. before elements the scheduler runs between
 * two members have updated their outputs this step.
 */
#ifdef USE_HIP

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../hines_ext.h"
#include "../hines_defs.h"
#include "hip_hsolve.h"
#include "hsolve_hip.h"

extern int do_chip_hh4_update(Hsolve *hsolve);   /* hines_4chip.c */

#define HIP_REFUSED ((void *)1)
#define HIP_GROUP_MIN 2          /* one-cell solvers worth merging */

struct hip_group;

typedef struct hip_state {
    Hsolve *hsolve;
    struct hsh_handle *h;        /* own handle; NULL while a group member */
    struct hip_group *grp;       /* the group it belongs to, or NULL */
    int member;                  /* its index in the group */
    hsh_member where;            /* its slices of the group's arrays */
    /* arrays handed to hsh_create (kept for a later group) */
    double *tab, *xv;
    int32_t nstab;
    /* message inputs: one hsh_input per ininfo entry, value read each step */
    hsh_input *in;
    Iinfo **in_src;
    int nin;
    /* spikegen: its SPIKE_OP and the messages GENESIS delivers (targets
     * that are not synchans of a device solver) */
    int32_t spike_op;
    MsgOut **ext;
    int next;
    /* connections to synchans of this solver: (SPIKE_OP, SYN2_OP, w, d) */
    int32_t *c_src, *c_dst;
    double *c_w, *c_d;
    int nconn;
    /* EVENTs posted by hip_synchan_event since the last step */
    int32_t *p_op, *p_T, *p_after;
    double *p_w;
    int npost, post_cap;
    double t_last;               /* SimulationTime() of the last step run */
    int32_t *spk_op, *spk_step;
    int spk_cap;
} HipState;

typedef struct hip_group {
    struct hsh_handle *h;
    HipState **m;
    int n;
    double t_last;
} HipGroup;

/* every solver the glue runs, in the order of its first PROCESS call */
static HipState **g_reg = NULL;
static int g_nreg = 0, g_capreg = 0;
static double g_reg_t = -1.0;    /* SimulationTime() when the last one joined */
static int g_grouped = 0;

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "HIP: out of host memory\n");
        exit(1);
    }
    return p;
}

/* The composite table and breakpoints gain one guard row (a copy of the top
 * row): the reference interpreter reads that row with weight 0 when
 * Vm >= xmax (hines_chip.c:181-186). */
static double *pad_rows(const double *src, int rows, int cols)
{
    double *p = (double *)xmalloc(sizeof(double) * (size_t)(rows + 1) * cols);
    memcpy(p, src, sizeof(double) * (size_t)rows * cols);
    memcpy(p + (size_t)rows * cols, src + (size_t)(rows - 1) * cols, sizeof(double) * cols);
    return p;
}

static void fill_model(Hsolve *hsolve, HipState *st, hsh_model *m)
{
    memset(m, 0, sizeof *m);
    m->ncompts = hsolve->ncompts;
    m->nchips = hsolve->nchips;
    m->nops = hsolve->nops;
    m->nfuncs = hsolve->nfuncs;
    m->nravals = hsolve->nravals;
    m->ncols = hsolve->ncols;
    m->xdivs = hsolve->xdivs;
    m->method = (BaseObject(hsolve)->method == CRANK_INT) ? HSH_CRANK : HSH_BEULER;
    m->chanmode = hsolve->chanmode;
    m->calcmode = hsolve->calcmode;
    m->dt = hsolve->dt;
    m->xmin = hsolve->xmin;
    m->invdx = hsolve->invdx;
    m->ops = hsolve->ops;
    m->chip = hsolve->chip;
    m->funcs = hsolve->funcs;
    m->ravals = hsolve->ravals;
    m->vm = hsolve->vm;
    /* pools, shells, buffers and pumps run on the device, in the program of
     * the compartment that owns them */
    m->nconcs = hsolve->nconcs;
    m->conc = hsolve->conc;
    m->nconcops = hsolve->concops ? hsolve->nconcops : 0;
    m->concops = hsolve->concops;
    m->nconcchip = hsolve->nconcchips;
    m->concchip = hsolve->concchip;
    m->ydivs = hsolve->ydivs;
    m->ymin = hsolve->ymin;
    m->invdy = hsolve->invdy;
    m->ngivals = hsolve->givals ? hsolve->ngivals : 0;
    m->givals = hsolve->givals;
    m->stablist = hsolve->stablist;
    m->nstab = st->nstab;
    m->tablist = st->tab;
    m->xvals = st->xv;
    /* diffusion systems (h_init_conc_solve, hines_conc_init.c:59-291) */
    if (hsolve->ndiffs > 0) {
        m->ndiffs = hsolve->ndiffs;
        m->nshells = hsolve->nshells;
        m->nshrows = hsolve->nshrows;
        m->nshfix = hsolve->nshfix;
        m->dconc = hsolve->dconc;
        m->fbufval = hsolve->fbufval;
        m->fdiffs = hsolve->fdiffs;
        m->ncovals0 = hsolve->fcoval[hsolve->ndiffs];
        m->covals0 = hsolve->covals0;
        m->nbufvals = hsolve->fbufval[hsolve->ndiffs];
        m->bufvals = hsolve->bufvals;
    }
}

/* Whether the solver's concops[] holds more than Ca_concen pools (shells,
 * pumps, buffers, concpools): such solvers are not merged into a group. */
static int conc_program(Hsolve *hsolve)
{
    int i;
    if (hsolve->ndiffs > 0) return 1;
    for (i = 0; hsolve->concops && i < hsolve->nconcops; i += 2) {
        if (hsolve->concops[i] == LCACONC_OP) break;
        if (hsolve->concops[i] != CACONC_OP) return 1;
    }
    return 0;
}

/* the ininfo list as device inputs (hines_msgs.c:329-348, :351-405) */
static void collect_inputs(Hsolve *hsolve, HipState *st)
{
    Iinfo *ii;
    int n = 0;
    for (ii = hsolve->ininfo; ii; ii = ii->next) n++;
    st->in = (hsh_input *)xmalloc(sizeof(hsh_input) * (size_t)n);
    st->in_src = (Iinfo **)xmalloc(sizeof(Iinfo *) * (size_t)n);
    st->nin = 0;
    for (ii = hsolve->ininfo; ii; ii = ii->next) {
        hsh_input *e = &st->in[st->nin];
        memset(e, 0, sizeof *e);
        e->index = ii->nchip;
        switch (ii->type) {
        case INJECT_OP: e->op = HSH_IN_INJECT; break;
        case BASE_OP: e->op = HSH_IN_BASE; break;
        case THRESH_OP: e->op = HSH_IN_THRESH; break;
        case FLUX_OP: e->op = HSH_IN_FLUX; break;
        case SYN2_OP:
            e->op = HSH_IN_ACTIVATION;
            e->stab = ii->nop;
            e->per_dt = (ii->msgin->type == ACTIVATION && ii->sChildType == SYNCHANC_T);
            break;
        default: continue;
        }
        st->in_src[st->nin++] = ii;
    }
}

/* Is dst a synchan this solver runs?  Its SYN2_OP index, or -1. */
static int32_t own_synchan(Hsolve *hsolve, Element *dst)
{
    int j;
    for (j = 0; j < hsolve->nchildren; j++)
        if (hsolve->children[j] == dst &&
            (hsolve->childtypes[j] == SYNCHAN_T || hsolve->childtypes[j] == SYNCHANC_T))
            return hsolve->childops[j];
    return -1;
}

/* The solver's spikegen: its SPIKE_OP, the messages to its own synchans as
 * device connections, the rest for GENESIS (h_dospike_event,
 * hines_conn.c:44-56).  A message to a synchan of another device solver
 * stays with GENESIS until the group is formed. */
static void collect_spikes(Hsolve *hsolve, HipState *st)
{
    MsgOut *mo;
    Element *sg = hsolve->spikegen;
    int j, n = 0;
    st->spike_op = -1;
    st->nconn = st->next = 0;
    if (!sg) return;
    for (j = 0; j < hsolve->nchildren; j++)
        if (hsolve->children[j] == sg && hsolve->childtypes[j] == SPIKEGEN_T) st->spike_op = hsolve->childops[j];
    MSGOUTLOOP(sg, mo) n++;
    st->ext = (MsgOut **)xmalloc(sizeof(MsgOut *) * (size_t)n);
    st->c_src = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)n);
    st->c_dst = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)n);
    st->c_w = (double *)xmalloc(sizeof(double) * (size_t)n);
    st->c_d = (double *)xmalloc(sizeof(double) * (size_t)n);
    MSGOUTLOOP(sg, mo) {
        int32_t op = own_synchan(hsolve, mo->dst);
        S2chan *sc = (S2chan *)mo->dst;
        int i, found = 0;
        if (op >= 0 && st->spike_op >= 0) {
            for (i = 0; i < sc->nsynapses; i++)
                if (sc->synapse[i].mi == (MsgIn *)mo) {
                    st->c_src[st->nconn] = st->spike_op;
                    st->c_dst[st->nconn] = op;
                    st->c_w[st->nconn] = sc->synapse[i].weight;
                    st->c_d[st->nconn] = sc->synapse[i].delay;
                    st->nconn++;
                    found = 1;
                    break;
                }
        }
        if (!found) st->ext[st->next++] = mo;
    }
}

/* Take st out of the registry, keeping the others' order (the order a
 * group's members take). */
static void reg_remove(HipState *st)
{
    int k, j = 0;
    for (k = 0; k < g_nreg; k++)
        if (g_reg[k] != st) g_reg[j++] = g_reg[k];
    g_nreg = j;
}

static void reg_add(HipState *st)
{
    if (g_nreg == g_capreg) {
        g_capreg = g_capreg ? 2 * g_capreg : 64;
        g_reg = (HipState **)realloc(g_reg, sizeof(HipState *) * (size_t)g_capreg);
        if (!g_reg) {
            fprintf(stderr, "HIP: out of host memory\n");
            exit(1);
        }
    }
    g_reg[g_nreg++] = st;
    g_reg_t = SimulationTime();
}

static int hip_init(Hsolve *hsolve)
{
    hsh_model m;
    hsh_options opt;
    HipState *st;
    int rc, i;

    if (hsolve->ydivs > hsolve->xdivs) {
        /* concentration rows beyond the composite table: CPU path */
        fprintf(stderr, "HIP: %s: concentration tables the device path does not run; "
                "CPU path for this solver\n", Pathname(hsolve));
        hsolve->accel_state = HIP_REFUSED;
        return -1;
    }
    st = (HipState *)calloc(1, sizeof(HipState));
    if (!st) return -1;
    st->hsolve = hsolve;
    st->nstab = 0;
    for (i = 0; i + 1 < hsolve->nops; i++)      /* synapse types referenced by ops[] */
        if ((hsolve->ops[i] == SYN2_OP || hsolve->ops[i] == SYN3_OP) && hsolve->ops[i + 1] / 4 + 1 > st->nstab)
            st->nstab = hsolve->ops[i + 1] / 4 + 1;
    if (hsolve->tablist && hsolve->ncols > 0) {
        st->tab = pad_rows(hsolve->tablist, hsolve->xdivs + 1, hsolve->ncols);
        st->xv = pad_rows(hsolve->xvals, hsolve->xdivs + 1, 1);
        st->xv[hsolve->xdivs + 1] = st->xv[hsolve->xdivs] + 1.0 / hsolve->invdx;
    }
    fill_model(hsolve, st, &m);
    memset(&opt, 0, sizeof opt);
    opt.flags = HSH_F_NO_AUTOTUNE;   /* one step per PROCESS call: no K to tune */
    rc = hsh_create(&m, &opt, &st->h);
    if (rc == HSH_REFUSED) {
        hsh_refusal r[16];
        int32_t n = 0;
        hsh_refusals(st->h, r, 16, &n);
        fprintf(stderr, "HIP: %s refused (%d compartment(s)); CPU path for this solver:\n",
                Pathname(hsolve), (int)n);
        for (i = 0; i < n && i < 16; i++)
            fprintf(stderr, "  compartment %d: code %d at index %d, reason %d\n",
                    r[i].comp, r[i].opcode, r[i].op_index, r[i].reason);
        hsh_destroy(st->h);
        free(st->tab);
        free(st->xv);
        free(st);
        hsolve->accel_state = HIP_REFUSED;
        return -1;
    }
    if (rc != HSH_OK) {
        fprintf(stderr, "HIP: %s: %s; CPU path for this solver\n", Pathname(hsolve), hsh_last_error());
        free(st->tab);
        free(st->xv);
        free(st);
        hsolve->accel_state = HIP_REFUSED;
        return -1;
    }
    /* events run on the device: the clock the lists are timed with, the
     * solver's own connections */
    hsh_set_time(st->h, SimulationTime());
    collect_inputs(hsolve, st);
    collect_spikes(hsolve, st);
    if (st->nconn > 0 &&
        hsh_set_synapses(st->h, st->nconn, st->c_src, st->c_dst, st->c_w, st->c_d) != HSH_OK) {
        fprintf(stderr, "HIP: %s: %s; CPU path for this solver\n", Pathname(hsolve), hsh_last_error());
        hsh_destroy(st->h);
        free(st);
        hsolve->accel_state = HIP_REFUSED;
        return -1;
    }
    st->spk_cap = 4096;
    st->spk_op = (int32_t *)xmalloc(sizeof(int32_t) * st->spk_cap);
    st->spk_step = (int32_t *)xmalloc(sizeof(int32_t) * st->spk_cap);
    st->t_last = -1.0;
    hsolve->accel_state = st;
    reg_add(st);
    return 0;
}

/* ---- group of one-cell solvers ----------------------------------------- */

static int group_eligible(HipState *st)
{
    return st->h && !st->grp && st->hsolve->ncompts == 1 && !conc_program(st->hsolve);
}

/* Merge the one-cell solvers that ran the first step into one handle, in
 * the order they ran it.  Their state moves over; connections between them
 * move onto the device. */
static void form_group(void)
{
    int i, j, k, n = 0, nconn = 0;
    HipGroup *g;
    hsh_model *ms;
    const hsh_model **mp;
    hsh_options opt;
    int32_t *cs, *cd;
    double *cw, *cdl;
    g_grouped = 1;
    for (i = 0; i < g_nreg; i++) if (group_eligible(g_reg[i])) n++;
    if (n < HIP_GROUP_MIN) return;
    g = (HipGroup *)calloc(1, sizeof(HipGroup));
    g->m = (HipState **)xmalloc(sizeof(HipState *) * (size_t)n);
    ms = (hsh_model *)xmalloc(sizeof(hsh_model) * (size_t)n);
    mp = (const hsh_model **)xmalloc(sizeof(hsh_model *) * (size_t)n);
    for (i = 0; i < g_nreg; i++) {
        HipState *st = g_reg[i];
        if (!group_eligible(st)) continue;
        /* the member's current state, in its host arrays */
        hip_sync_chip(st->hsolve);
        fill_model(st->hsolve, st, &ms[g->n]);
        mp[g->n] = &ms[g->n];
        g->m[g->n++] = st;
    }
    memset(&opt, 0, sizeof opt);
    opt.flags = HSH_F_NO_AUTOTUNE;
    /* a synchan with an event pending after the first step (its counter in
     * ops[] >= 0) makes the group refuse (HSH_R_SYN_PENDING): the solvers
     * then keep their own handles and lists */
    if (hsh_create_group(mp, g->n, &opt, &g->h) != HSH_OK) {
        fprintf(stderr, "HIP: one-cell solvers stay separate: %s\n", hsh_last_error());
        free(g->m);
        free(g);
        free(ms);
        free(mp);
        return;
    }
    hsh_set_time(g->h, SimulationTime());      /* the group runs this step next */
    for (i = 0; i < g->n; i++) {
        HipState *st = g->m[i];
        st->grp = g;
        st->member = i;
        hsh_group_member(g->h, i, &st->where);
        nconn += st->nconn + st->next;
    }
    /* device connections: each member's own, and its spikegen's messages to
     * synchans of other members (in group ops indices) */
    cs = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)nconn);
    cd = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)nconn);
    cw = (double *)xmalloc(sizeof(double) * (size_t)nconn);
    cdl = (double *)xmalloc(sizeof(double) * (size_t)nconn);
    k = 0;
    for (i = 0; i < g->n; i++) {
        HipState *st = g->m[i];
        int keep = 0;
        for (j = 0; j < st->nconn; j++) {
            cs[k] = st->c_src[j] + st->where.op0;
            cd[k] = st->c_dst[j] + st->where.op0;
            cw[k] = st->c_w[j];
            cdl[k] = st->c_d[j];
            k++;
        }
        for (j = 0; j < st->next; j++) {
            MsgOut *mo = st->ext[j];
            Element *dst = mo->dst;
            HipState *ds = NULL;
            int32_t op = -1;
            int q;
            if (IsHsolved(dst)) {
                S2chan *sc = (S2chan *)dst;
                Hsolve *dh = (Hsolve *)sc->hsolve;
                if (dh && dh->accel_state && dh->accel_state != HIP_REFUSED) ds = (HipState *)dh->accel_state;
                if (ds && ds->grp == g) op = own_synchan(dh, dst);
                if (op >= 0) {
                    for (q = 0; q < sc->nsynapses; q++)
                        if (sc->synapse[q].mi == (MsgIn *)mo) break;
                    if (q < sc->nsynapses && st->spike_op >= 0) {
                        cs[k] = st->spike_op + st->where.op0;
                        cd[k] = op + ds->where.op0;
                        cw[k] = sc->synapse[q].weight;
                        cdl[k] = sc->synapse[q].delay;
                        k++;
                        continue;
                    }
                }
            }
            st->ext[keep++] = mo;        /* still delivered by GENESIS */
        }
        st->next = keep;
    }
    if (k > 0 && hsh_set_synapses(g->h, k, cs, cd, cw, cdl) != HSH_OK)
        fprintf(stderr, "HIP: group connections: %s\n", hsh_last_error());
    free(cs);
    free(cd);
    free(cw);
    free(cdl);
    free(ms);
    free(mp);
    for (i = 0; i < g->n; i++) {
        hsh_destroy(g->m[i]->h);
        g->m[i]->h = NULL;
    }
    g->t_last = -1.0;
}

/* ---- per step ------------------------------------------------------------- */

static void push_inputs(struct hsh_handle *h, HipState *st, int32_t chip0)
{
    int i;
    if (st->nin == 0) return;
    for (i = 0; i < st->nin; i++) {
        Iinfo *ii = st->in_src[i];
        st->in[i].value = MSGVALUE(ii->msgin, 0);
    }
    if (chip0 == 0) {
        if (hsh_set_inputs_static(h, st->nin, st->in) != HSH_OK)
            fprintf(stderr, "HIP: inputs: %s\n", hsh_last_error());
    } else {
        /* a member: its chip indices are offset in the group */
        hsh_input tmp[64];
        int done = 0;
        while (done < st->nin) {
            int n = st->nin - done < 64 ? st->nin - done : 64;
            for (i = 0; i < n; i++) {
                tmp[i] = st->in[done + i];
                if (tmp[i].op != HSH_IN_FLUX) tmp[i].index += chip0;
                else tmp[i].index += st->where.conc0;
                if (tmp[i].op == HSH_IN_ACTIVATION) tmp[i].stab += st->where.stab0;
            }
            if (hsh_set_inputs_static(h, n, tmp) != HSH_OK)
                fprintf(stderr, "HIP: inputs: %s\n", hsh_last_error());
            done += n;
        }
    }
}

static void flush_posts(struct hsh_handle *h, HipState *st, int32_t op0)
{
    int i, a;
    if (st->npost == 0) return;
    for (i = 0; i < st->npost; i++) st->p_op[i] += op0;
    /* events of the last step's end first, then those of this step's start
     * (hsh_post_events keeps the order within each) */
    for (a = 1; a >= 0; a--) {
        int n = 0;
        int32_t *op = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)st->npost);
        int32_t *T = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)st->npost);
        double *w = (double *)xmalloc(sizeof(double) * (size_t)st->npost);
        for (i = 0; i < st->npost; i++)
            if (st->p_after[i] == a) {
                op[n] = st->p_op[i];
                T[n] = st->p_T[i];
                w[n] = st->p_w[i];
                n++;
            }
        if (n > 0 && hsh_post_events(h, n, op, T, w, a) != HSH_OK)
            fprintf(stderr, "HIP: events: %s\n", hsh_last_error());
        free(op);
        free(T);
        free(w);
    }
    st->npost = 0;
}

/* spikes of this solver to elements GENESIS reaches: only those messages */
static void deliver_spikes(Hsolve *hsolve, HipState *st, int32_t nspk)
{
    Spike *spike = (Spike *)hsolve->spikegen;
    int i, j;
    if (!spike) return;
    for (i = 0; i < nspk; i++) {
        for (j = 0; j < st->next; j++) CallEventAction(spike, st->ext[j]);
        spike->lastevent = SimulationTime();
    }
}

int hip_chip_update(Hsolve *hsolve)
{
    HipState *st;
    hsh_step_out o;
    int32_t n = 0;
    int i;
    double t = SimulationTime();

    if (hsolve->accel_state == NULL && hip_init(hsolve) != 0) {
        /* fall through to the CPU path below */
    }
    if (hsolve->accel_state == HIP_REFUSED) {
        do_chip_hh4_update(hsolve);
        return 0;                       /* hines.c runs the CPU elimination */
    }
    st = (HipState *)hsolve->accel_state;
    /* the first PROCESS of a later step: the solvers that ran the previous
     * step are known, merge the one-cell ones */
    if (!g_grouped && t > g_reg_t) form_group();
    if (st->grp) {
        HipGroup *g = st->grp;
        if (g->t_last != t) {
            /* the first member's turn this step: inputs and events of every
             * member, one launch, spikes to GENESIS */
            for (i = 0; i < g->n; i++) {
                push_inputs(g->h, g->m[i], g->m[i]->where.chip0);
                flush_posts(g->h, g->m[i], g->m[i]->where.op0);
            }
            if (hsh_step(g->h, 1, HSH_BATCHED, &o) != HSH_OK) {
                fprintf(stderr, "HIP: step failed: %s\n", hsh_last_error());
                return -1;
            }
            g->t_last = t;
            for (i = 0; i < g->n; i++) g->m[i]->t_last = t;
            if (hsh_drain_spikes(g->h, st->spk_op, st->spk_step, st->spk_cap, &n) == HSH_OK) {
                int k;
                for (k = 0; k < n; k++)
                    for (i = 0; i < g->n; i++)
                        if (g->m[i]->spike_op + g->m[i]->where.op0 == st->spk_op[k])
                            deliver_spikes(g->m[i]->hsolve, g->m[i], 1);
            }
        }
        if (hsolve->outinfo) hip_sync_chip(hsolve);
        return 1;
    }
    push_inputs(st->h, st, 0);
    flush_posts(st->h, st, 0);
    if (hsh_step(st->h, 1, HSH_BATCHED, &o) != HSH_OK) {
        fprintf(stderr, "HIP: step failed: %s\n", hsh_last_error());
        return -1;
    }
    st->t_last = t;
    if (hsolve->spikegen && st->next > 0) {
        if (hsh_drain_spikes(st->h, st->spk_op, st->spk_step, st->spk_cap, &n) == HSH_OK)
            deliver_spikes(hsolve, st, n);
    } else if (hsolve->spikegen) {
        hsh_drain_spikes(st->h, st->spk_op, st->spk_step, st->spk_cap, &n);
    }
    if (hsolve->outinfo) hsh_get_state(st->h, hsolve->vm, NULL, NULL, NULL);
    return 1;
}

/* ---- EVENTs from outside --------------------------------------------------- */

int hip_synchan_event(Element *hsolve_elm, int solve_index, int time, double weight)
{
    Hsolve *hsolve = (Hsolve *)hsolve_elm;
    HipState *st;
    if (!hsolve || !hsolve->accel_state || hsolve->accel_state == HIP_REFUSED) return 0;
    st = (HipState *)hsolve->accel_state;
    if (st->npost == st->post_cap) {
        st->post_cap = st->post_cap ? 2 * st->post_cap : 256;
        st->p_op = (int32_t *)realloc(st->p_op, sizeof(int32_t) * (size_t)st->post_cap);
        st->p_T = (int32_t *)realloc(st->p_T, sizeof(int32_t) * (size_t)st->post_cap);
        st->p_after = (int32_t *)realloc(st->p_after, sizeof(int32_t) * (size_t)st->post_cap);
        st->p_w = (double *)realloc(st->p_w, sizeof(double) * (size_t)st->post_cap);
        if (!st->p_op || !st->p_T || !st->p_after || !st->p_w) {
            fprintf(stderr, "HIP: out of host memory\n");
            exit(1);
        }
    }
    st->p_op[st->npost] = hsolve->childops[solve_index];
    st->p_T[st->npost] = time;
    st->p_w[st->npost] = weight;
    /* the solver has run this step already: the event comes after its ops */
    st->p_after[st->npost] = (st->t_last == SimulationTime());
    st->npost++;
    return 1;
}

/* ---- state on demand -------------------------------------------------------- */

void hip_sync_chip(Hsolve *hsolve)
{
    HipState *st = (HipState *)hsolve->accel_state;
    if (!st || st == (HipState *)HIP_REFUSED) return;
    if (st->grp) {
        /* the member's slices of the group's arrays */
        struct hsh_handle *h = st->grp->h;
        hsh_info info;
        double *vm, *chip, *res, *conc = NULL, *giv = NULL;
        int32_t *ops;
        int32_t nrows = 0, nchips = 0, nops = 0, nconcs = 0, ngiv = 0, i;
        hsh_member e;
        HipGroup *g = st->grp;
        for (i = 0; i < g->n; i++) {
            hsh_group_member(h, i, &e);
            nrows += e.nrows;
            nchips += e.nchips;
            nops += e.nops;
            nconcs += e.nconcs;
            ngiv += e.ngivals;
        }
        (void)info;
        vm = (double *)xmalloc(sizeof(double) * (size_t)nrows);
        res = (double *)xmalloc(sizeof(double) * 2 * (size_t)nrows);
        chip = (double *)xmalloc(sizeof(double) * (size_t)nchips);
        ops = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)(nops + 1));
        if (hsh_get_state(h, vm, chip, res, ops) == HSH_OK) {
            e = st->where;
            memcpy(hsolve->vm, vm + e.row0, sizeof(double) * (size_t)e.nrows);
            memcpy(hsolve->results, res + 2 * e.row0, sizeof(double) * 2 * (size_t)e.nrows);
            memcpy(hsolve->chip, chip + e.chip0, sizeof(double) * (size_t)e.nchips);
            memcpy(hsolve->ops, ops + e.op0, sizeof(int32_t) * (size_t)e.nops);
            if (nconcs > 0 && e.nconcs > 0) {
                conc = (double *)xmalloc(sizeof(double) * (size_t)nconcs);
                if (hsh_get_conc(h, conc) == HSH_OK)
                    memcpy(hsolve->conc, conc + e.conc0, sizeof(double) * (size_t)e.nconcs);
            }
            if (ngiv > 0 && e.ngivals > 0 && hsolve->givals) {
                giv = (double *)xmalloc(sizeof(double) * (size_t)ngiv);
                if (hsh_get_givals(h, giv) == HSH_OK)
                    memcpy(hsolve->givals, giv + e.giv0, sizeof(double) * (size_t)e.ngivals);
            }
        }
        free(vm);
        free(res);
        free(chip);
        free(ops);
        free(conc);
        free(giv);
        return;
    }
    hsh_get_state(st->h, hsolve->vm, hsolve->chip, hsolve->results, hsolve->ops);
    if (hsolve->nconcs > 0) hsh_get_conc(st->h, hsolve->conc);
    if (hsolve->givals && hsolve->ngivals > 0) hsh_get_givals(st->h, hsolve->givals);
}

/* After HPUT, HCHANGE, HRESTORE or RESTORE2 changed the host arrays: the
 * device takes them (hines.c:476-571). */
void hip_push_state(Hsolve *hsolve)
{
    HipState *st = (HipState *)hsolve->accel_state;
    if (!st || st == (HipState *)HIP_REFUSED) return;
    if (st->grp) {
        /* the group's arrays with this member's slices replaced */
        struct hsh_handle *h = st->grp->h;
        HipGroup *g = st->grp;
        hsh_member e;
        double *vm, *chip, *conc = NULL;
        int32_t *ops;
        int32_t nrows = 0, nchips = 0, nops = 0, nconcs = 0, i;
        for (i = 0; i < g->n; i++) {
            hsh_group_member(h, i, &e);
            nrows += e.nrows;
            nchips += e.nchips;
            nops += e.nops;
            nconcs += e.nconcs;
        }
        vm = (double *)xmalloc(sizeof(double) * (size_t)nrows);
        chip = (double *)xmalloc(sizeof(double) * (size_t)nchips);
        ops = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)(nops + 1));
        if (hsh_get_state(h, vm, chip, NULL, ops) == HSH_OK) {
            e = st->where;
            memcpy(vm + e.row0, hsolve->vm, sizeof(double) * (size_t)e.nrows);
            memcpy(chip + e.chip0, hsolve->chip, sizeof(double) * (size_t)e.nchips);
            memcpy(ops + e.op0, hsolve->ops, sizeof(int32_t) * (size_t)e.nops);
            hsh_put_state(h, vm, chip, ops);
            if (nconcs > 0 && e.nconcs > 0) {
                conc = (double *)xmalloc(sizeof(double) * (size_t)nconcs);
                if (hsh_get_conc(h, conc) == HSH_OK) {
                    memcpy(conc + e.conc0, hsolve->conc, sizeof(double) * (size_t)e.nconcs);
                    hsh_put_conc(h, conc);
                }
            }
        }
        free(vm);
        free(chip);
        free(ops);
        free(conc);
        return;
    }
    hsh_put_state(st->h, hsolve->vm, hsolve->chip, hsolve->ops);
    if (hsolve->nconcs > 0) hsh_put_conc(st->h, hsolve->conc);
}

int hip_owns_conc(Hsolve *hsolve)
{
    HipState *st = (HipState *)hsolve->accel_state;
    return st != NULL && st != (HipState *)HIP_REFUSED;
}

static void free_state(HipState *st)
{
    free(st->tab);
    free(st->xv);
    free(st->in);
    free(st->in_src);
    free(st->ext);
    free(st->c_src);
    free(st->c_dst);
    free(st->c_w);
    free(st->c_d);
    free(st->p_op);
    free(st->p_T);
    free(st->p_after);
    free(st->p_w);
    free(st->spk_op);
    free(st->spk_step);
    free(st);
}

/* h_delete, RESET: this solver's device state goes; a group it belongs to
 * is dissolved and its members start afresh at their next PROCESS (after a
 * RESET every solver's state is new anyway). */
void hip_release(Hsolve *hsolve)
{
    HipState *st = (HipState *)hsolve->accel_state;
    int i;
    if (!st || st == (HipState *)HIP_REFUSED) {
        hsolve->accel_state = NULL;
        return;
    }
    if (st->grp) {
        HipGroup *g = st->grp;
        hsh_destroy(g->h);
        for (i = 0; i < g->n; i++) {
            HipState *m = g->m[i];
            m->hsolve->accel_state = NULL;
            reg_remove(m);
            free_state(m);
        }
        free(g->m);
        free(g);
    } else {
        hsh_destroy(st->h);
        reg_remove(st);
        free_state(st);
        hsolve->accel_state = NULL;
    }
    g_grouped = 0;
}

#endif /* USE_HIP */
