/*
 * hsolve_oracle.h -- CPU restatement of GENESIS's chanmode-4/5 hsolve step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle: tests/, the smoke
 * check and the CPU-baseline leg of bench.py call it to produce the reference
 * answer.  No product code path links, loads or calls it.
 *
 * It walks the same Hsolve-format arrays as the reference CPU code, in the
 * same array-of-structures order:
 *   ohs_chip_update   genesis/src/hines/hines_chip.c:78-602 (CHANMODE4 with
 *                     SOLVEINTERPOL, i.e. do_chip_hh4_update, calcmode 1)
 *   ohs_euler_solve   genesis/src/hines/hines_solve.c:223-296
 *   ohs_crank_solve   genesis/src/hines/hines_solve.c:140-218
 *   ohs_step          phase order of genesis/src/hines/hines.c:222-270
 * Compile with -O2 -ffp-contract=off and no fast-math.
 */
#ifndef HSOLVE_ORACLE_H
#define HSOLVE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ohs ohs;

/* synchan event hook: called when a SYN2/SYN3 counter reaches 0
 * (h_dosynchan, hines_conn.c:60-89).  It may add to chip[xslot] and must
 * return the next event count (or -1 for none). */
typedef int32_t (*ohs_syn_fn)(void *ctx, int32_t stabindex, int32_t child,
                              int32_t xslot, int32_t step);
/* uniform random number for SYN3_OP (urandom, hines_chip.c:516) */
typedef double (*ohs_rand_fn)(void *ctx);

typedef struct ohs_arrays {
    int32_t ncompts, nchips, nops, nfuncs, nravals;
    int32_t ncols, xdivs, ydivs, nconcs, nstab, ngivals, nelm_names;
    int32_t method, chanmode, storemode, pad;
    double  xmin, invdx, ymin, invdy, dt;
    const int32_t *ops;
    const double  *chip;
    const double  *tablist;   /* (xdivs+2)*ncols, guard row included */
    const double  *xvals;     /* xdivs+2 */
    const double  *yvals;     /* ydivs+2 or NULL */
    const double  *stablist;
    const int32_t *funcs;
    const double  *ravals;
    const double  *vm;
    const double  *givals;    /* may be NULL: then no givals are kept */
    const double  *conc;      /* nconcs or NULL */
    const int32_t *concops;   /* nconcops or NULL: the concentration program
                                 (hines_conc_init.c:426-805): CONC_OP / FCONC_OP
                                 index, FLUXCONC_OP, the pumps, BUFFER_OP, POOL_OP,
                                 CACONC_OP, ended by LCONC_OP or LCACONC_OP */
    const double  *concchip;  /* nconcchip or NULL: the values those ops consume */
    int32_t nconcops, nconcchip;
    /* Diffusion systems of shells and buffers (do_h_conc_solve,
     * hines_conc.c:39-396; arrays of h_init_conc_solve,
     * hines_conc_init.c:59-291).  System i has nshells[i] shells of
     * nshrows[i] rows each (its buffers, the nshfix[i] fixed ones first, then
     * the shell); its concentrations start at conc[dconc[i]], row r of shell m
     * at dconc[i] + m*nshrows[i] + r; its buffer rates at bufvals[fbufval[i]]
     * (3 per buffer).  covals0 holds every system's coupling and buffer-total
     * values in the order the solve consumes them.  ndiffs = 0: none. */
    int32_t ndiffs, ncovals0, nbufvals, pad2;
    const int32_t *nshells, *nshrows, *nshfix, *dconc, *fbufval;
    const double  *covals0, *bufvals;
} ohs_arrays;

ohs    *ohs_create(const ohs_arrays *a);
void    ohs_destroy(ohs *h);
void    ohs_set_hooks(ohs *h, ohs_syn_fn syn, ohs_rand_fn rnd, void *ctx);

/* One interpreter pass (do_chip_hh4_update).  Returns 0 or -1 on an
 * undefined opcode. */
int32_t ohs_chip_update(ohs *h);
void    ohs_euler_solve(ohs *h);
void    ohs_crank_solve(ohs *h);
/* The pass over concops[] (do_h4_conc_chip_update): shell results and
 * diagonals, pumps, pool buffers, pools and Ca_concen.  Returns 0 or -1 on an
 * undefined opcode. */
int32_t ohs_conc_chip_update(ohs *h);
/* The diffusion systems' implicit step (do_h_conc_solve).  Returns 0, or -1
 * for a system of fewer than two shells. */
int32_t ohs_conc_solve(ohs *h);
/* codiag[] and coresult[] (nconcs each; either may be NULL): the per
 * concentration diagonal and result the concops pass leaves for the solve */
void    ohs_conc_regs(ohs *h, double *codiag, double *coresult);
void    ohs_get_conc(ohs *h, double *conc);
void    ohs_put_conc(ohs *h, const double *conc);
/* nsteps full steps: chip update, the concops pass, the diffusion solve, then
 * the elimination selected by method. */
int32_t ohs_step(ohs *h, int32_t nsteps);

/* state access (copies) */
void    ohs_get(ohs *h, double *vm, double *chip, double *results, int32_t *ops,
                double *givals);
void    ohs_put(ohs *h, const double *vm, const double *chip, const int32_t *ops);

/* spike events raised by SPIKE_OP: ops index of the SPIKE_OP and step. */
int32_t ohs_nspikes(ohs *h);
void    ohs_spikes(ohs *h, int32_t *opidx, int32_t *step);
void    ohs_clear_spikes(ohs *h);
int32_t ohs_stepno(ohs *h);

/* Per-compartment operation trace of the last ohs_chip_update: for each
 * compartment, a 64-bit FNV-1a hash over the opcodes and operands it
 * executed in order.  Used to check that an ops walk visits the same ops. */
void    ohs_trace_enable(ohs *h, int32_t on);
void    ohs_trace(ohs *h, uint64_t *per_comp);

/* The gate-table row (filo) each compartment's NEWVOLT_OP used in the last
 * ohs_chip_update (-1: no NEWVOLT), recorded once enabled.  Two runs whose
 * rows agree at every step interpolate the same table intervals: their
 * difference quotient measures the derivative of one piecewise-linear
 * model. */
void    ohs_rows_enable(ohs *h, int32_t on);
void    ohs_rows(ohs *h, int32_t *rows);

/* Step nh independent solvers (disjoint neuron blocks, each its own hsolve)
 * nsteps times on nthreads POSIX threads; solver i runs on thread i % nthreads.
 * Used for the multi-core CPU baseline. */
int32_t ohs_step_parallel(ohs **hs, int32_t nh, int32_t nsteps, int32_t nthreads);

/* Synaptic connections and the synchans' pending event lists, as hsolved
 * synchans keep them (newconn/synchan.c:161-358, hines.c:352-357,
 * hines_conn.c:44-89).  A spike of the SPIKE_OP at ops index src_op[i] is an
 * EVENT at the synchan whose SYN2_OP is at dst_op[i], in connection order:
 * it hits time = (int)((float)delay[i]/dt + 0.9) steps later; a non-empty
 * list first re-times its head by (int)((t - time_last_event)/dt) of the
 * floating simulation times, the event goes before the first one whose time
 * is >= its own (times relative to the event before), and a new head sets
 * the SYN2_OP counter.  When the counter reaches 0 the synchan pops the
 * front events whose time is 0, sums (float)weight/dt into its float
 * activation, adds dt*activation*stablist[k+3] to X and sets the counter to
 * the new head's time or -1.  Returns the shortest connection in steps
 * (`time` of the shortest delay), or -1 on bad arguments.  The pending
 * lists persist across calls; n = 0 removes the connections. */
int32_t ohs_set_synapses(ohs *h, int32_t n, const int32_t *src_op, const int32_t *dst_op,
                         const double *weight, const double *delay);
/* An EVENT from outside the solver: `time` steps to the hit (the caller's
 * (int)((delay - (t - eventtime))/dt + 0.9)), a float weight.  after = 0: at
 * the simulation time of the next step, before its ops; after = 1: at the
 * time of the last step run, after its ops.  Returns 0 or -1. */
int32_t ohs_post_event(ohs *h, int32_t syn_op, int32_t time, double weight, int32_t after);
/* One message input of h_in_msgs (hines_msgs.c:351-405): op is INJECT_OP
 * (1), BASE_OP (2), THRESH_OP (3) -- chip[index] = value -- SYN2_OP (4100)
 * -- an ACTIVATION to the synchan whose X is chip[index]: when value >
 * VTINY, chip[index] += value*stablist[stab+3], times dt for a synchan
 * (not a channelC2/C3) -- or FLUX_OP (3300): flux[index] = value. */
typedef struct ohs_input {
    int32_t op, index, stab, per_dt;
    double value;
} ohs_input;
/* Apply inputs in order, as h_in_msgs walks its ininfo list before the
 * chip update.  Returns 0 or -1 on an unknown op or index. */
int32_t ohs_in_msgs(ohs *h, int32_t n, const ohs_input *in);

/* SimulationTime() of the next step (0 at creation); it grows by dt per
 * step (simulation_time += ClockValue(0), sim/sim_step.c:112). */
void    ohs_set_time(ohs *h, double t);
double  ohs_time(ohs *h);
/* The pending list of the synchan at ops index syn_op: its events' time
 * fields and weight/dt (up to cap), time_last_event; returns the list
 * length, -1 if the synchan has none. */
int32_t ohs_pending(ohs *h, int32_t syn_op, int32_t *time, double *wdt, int32_t cap, double *tle);

/* Tangent-linear model (hsolve_tl.c): the same step carrying the derivative
 * of every value along one direction of (vm0, chip, ravals).  Tabchannel
 * subset only; ohs_tl_create returns NULL for other ops. */
typedef struct ohs_tl ohs_tl;
ohs_tl *ohs_tl_create(const ohs_arrays *a);
void    ohs_tl_destroy(ohs_tl *t);
/* tangents of vm[ncompts], chip[nchips], ravals[nravals] (NULL = zero) */
void    ohs_tl_set_tangent(ohs_tl *t, const double *dvm, const double *dchip,
                           const double *dravals);
/* nsteps steps; after step s, trace[s*nobs+j] = vm[obs[j]] and dtrace the
 * same for its tangent (either may be NULL) */
int32_t ohs_tl_step(ohs_tl *t, int32_t nsteps, const int32_t *obs, int32_t nobs,
                    double *trace, double *dtrace);
void    ohs_tl_get(ohs_tl *t, double *vm, double *dvm, double *chip, double *dchip);

#ifdef __cplusplus
}
#endif

#endif
