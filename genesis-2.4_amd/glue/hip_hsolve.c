/*
 * hip_hsolve.c -- GENESIS glue for libhsolve_hip (compiled inside a GENESIS
 * tree, genesis/src/hines/hip/, with -DUSE_HIP; it includes GENESIS headers
 * and is not part of libhsolve_hip).
 *
 * Mirrors the control flow of the CUDA glue (genesis/src/hines/cuda/
 * cuda_hsolve.c:289-517): lazy creation on the first PROCESS call, refusal
 * -> CPU path for this solver with the opcode list on stderr, per-step
 * device step, state download on demand.  The device keeps the state between
 * steps; only the message-driven chip values are pushed each step and vm is
 * pulled only when outputs need it.
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
extern void h_dospike_event(Hsolve *hsolve);     /* hines_conn.c:44 */

#define HIP_REFUSED ((void *)1)

typedef struct hip_state {
    struct hsh_handle *h;
    int32_t *spk_op, *spk_step;
    int spk_cap;
} HipState;

/* The composite table and breakpoints gain one guard row (a copy of the top
 * row): the reference interpreter reads that row with weight 0 when
 * Vm >= xmax (hines_chip.c:181-186). */
static double *pad_rows(const double *src, int rows, int cols)
{
    double *p = (double *)malloc(sizeof(double) * (size_t)(rows + 1) * cols);
    if (!p) return NULL;
    memcpy(p, src, sizeof(double) * (size_t)rows * cols);
    memcpy(p + (size_t)rows * cols, src + (size_t)(rows - 1) * cols, sizeof(double) * cols);
    return p;
}

static int hip_init(Hsolve *hsolve)
{
    hsh_model m;
    hsh_options opt;
    HipState *st;
    double *tab = NULL, *xv = NULL;
    int rc, i;

    memset(&m, 0, sizeof m);
    m.ncompts = hsolve->ncompts;
    m.nchips = hsolve->nchips;
    m.nops = hsolve->nops;
    m.nfuncs = hsolve->nfuncs;
    m.nravals = hsolve->nravals;
    m.ncols = hsolve->ncols;
    m.xdivs = hsolve->xdivs;
    m.method = (BaseObject(hsolve)->method == CRANK_INT) ? HSH_CRANK : HSH_BEULER;
    m.chanmode = hsolve->chanmode;
    m.calcmode = hsolve->calcmode;
    m.dt = hsolve->dt;
    m.xmin = hsolve->xmin;
    m.invdx = hsolve->invdx;
    m.ops = hsolve->ops;
    m.chip = hsolve->chip;
    m.funcs = hsolve->funcs;
    m.ravals = hsolve->ravals;
    m.vm = hsolve->vm;
    /* concentrations: Ca_concen pools run on the device, in the program of
     * the compartment that feeds them */
    m.nconcs = hsolve->nconcs;
    m.conc = hsolve->conc;
    m.nconcops = hsolve->concops ? hsolve->nconcops : 0;
    m.concops = hsolve->concops;
    m.nconcchip = hsolve->nconcchips;
    m.concchip = hsolve->concchip;
    m.ydivs = hsolve->ydivs;
    m.ymin = hsolve->ymin;
    m.invdy = hsolve->invdy;
    /* chanmode-4/5 output stream: written on the device, pulled on demand */
    m.ngivals = hsolve->givals ? hsolve->ngivals : 0;
    m.givals = hsolve->givals;
    if (hsolve->ydivs > hsolve->xdivs || hsolve->ndiffs) {
        /* concentration rows beyond the composite table, or a conc solve
         * (difshells): CPU path */
        fprintf(stderr, "HIP: %s: concentration tables or shells the device path does not run; "
                "CPU path for this solver\n", Pathname(hsolve));
        hsolve->accel_state = HIP_REFUSED;
        return -1;
    }
    m.stablist = hsolve->stablist;
    m.nstab = 0;
    for (i = 0; i + 1 < hsolve->nops; i++)      /* synapse types referenced by ops[] */
        if ((hsolve->ops[i] == SYN2_OP || hsolve->ops[i] == SYN3_OP) &&
            hsolve->ops[i + 1] / 4 + 1 > m.nstab)
            m.nstab = hsolve->ops[i + 1] / 4 + 1;
    if (hsolve->tablist && hsolve->ncols > 0) {
        tab = pad_rows(hsolve->tablist, hsolve->xdivs + 1, hsolve->ncols);
        xv = pad_rows(hsolve->xvals, hsolve->xdivs + 1, 1);
        xv[hsolve->xdivs + 1] = xv[hsolve->xdivs] + 1.0 / hsolve->invdx;
        m.tablist = tab;
        m.xvals = xv;
    }
    memset(&opt, 0, sizeof opt);
    /* one step per PROCESS call: no K to tune, default layout */
    opt.flags = HSH_F_NO_AUTOTUNE;
    st = (HipState *)calloc(1, sizeof(HipState));
    rc = hsh_create(&m, &opt, &st->h);
    free(tab);
    free(xv);
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
        free(st);
        hsolve->accel_state = HIP_REFUSED;
        return -1;
    }
    if (rc != HSH_OK) {
        fprintf(stderr, "HIP: %s: %s; CPU path for this solver\n", Pathname(hsolve), hsh_last_error());
        free(st);
        hsolve->accel_state = HIP_REFUSED;
        return -1;
    }
    st->spk_cap = 4096;
    st->spk_op = (int32_t *)malloc(sizeof(int32_t) * st->spk_cap);
    st->spk_step = (int32_t *)malloc(sizeof(int32_t) * st->spk_cap);
    hsolve->accel_state = st;
    return 0;
}

int hip_chip_update(Hsolve *hsolve)
{
    HipState *st;
    hsh_step_out o;
    int32_t n = 0;
    int i;

    if (hsolve->accel_state == NULL && hip_init(hsolve) != 0) {
        /* fall through to the CPU path below */
    }
    if (hsolve->accel_state == HIP_REFUSED) {
        do_chip_hh4_update(hsolve);
        return 0;                       /* hines.c runs the CPU elimination */
    }
    st = (HipState *)hsolve->accel_state;
    /* h_in_msgs (hines_msgs.c:351-405) has just written message values into
     * chip[]; a GENESIS build would keep the list of chip indices its ininfo
     * entries target and push them with hsh_set_chip here. */
    /* one step, channels and elimination in one fused launch (the batched
     * kernel with K = 1; HSH_PLAIN would make two launches) */
    if (hsh_step(st->h, 1, HSH_BATCHED, &o) != HSH_OK) {
        fprintf(stderr, "HIP: step failed: %s\n", hsh_last_error());
        return -1;
    }
    if (hsolve->spikegen) {
        if (hsh_drain_spikes(st->h, st->spk_op, st->spk_step, st->spk_cap, &n) == HSH_OK)
            for (i = 0; i < n; i++) h_dospike_event(hsolve);
    }
    if (hsolve->outinfo) hsh_get_state(st->h, hsolve->vm, NULL, NULL, NULL);
    return 1;
}

void hip_sync_chip(Hsolve *hsolve)
{
    HipState *st = (HipState *)hsolve->accel_state;
    if (!st || st == (HipState *)HIP_REFUSED) return;
    hsh_get_state(st->h, hsolve->vm, hsolve->chip, hsolve->results, hsolve->ops);
    if (hsolve->nconcs > 0) hsh_get_conc(st->h, hsolve->conc);
    if (hsolve->givals && hsolve->ngivals > 0) hsh_get_givals(st->h, hsolve->givals);
}

int hip_owns_conc(Hsolve *hsolve)
{
    HipState *st = (HipState *)hsolve->accel_state;
    return st != NULL && st != (HipState *)HIP_REFUSED;
}

void hip_release(Hsolve *hsolve)
{
    HipState *st = (HipState *)hsolve->accel_state;
    if (!st || st == (HipState *)HIP_REFUSED) return;
    hsh_destroy(st->h);
    free(st->spk_op);
    free(st->spk_step);
    free(st);
    hsolve->accel_state = NULL;
}

#endif /* USE_HIP */
