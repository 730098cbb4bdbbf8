/*
 * hsm_model.h -- setup compiler for synthetic hsolve populations (CPU, C ABI).
 *
 * GENESIS builds the arrays that the chanmode-4/5 hsolve step consumes at
 * SETUP/RESET time: the Hines ordering and elimination program
 * (genesis/src/hines/hines_init.c:40-345, :361-1273), the channel program and
 * its chip[] layout (hines_chip_init.c:465-1260), the composite gate table
 * (hines_chip_init.c:1262-1400, olf/new_interp.c:533-583) and the RESET values
 * (hines_chip_init.c:1402-1552, hines_data.c:80-535).  No GENESIS runtime is
 * available here, so this library restates those setup steps for populations
 * described by plain arrays, and emits the same array formats.  Its output is
 * the input of both the device library (libhsolve_hip) and the CPU oracle.
 *
 * Everything is fp64.  Units are SI (V, F, Ohm, A, S, s).
 */
#ifndef HSM_MODEL_H
#define HSM_MODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSM_BEULER 10
#define HSM_CRANK  11

/* One voltage-dependent gate of a tabchannel, as set up by `setupalpha`
 * (olf/new_interp.c:533-583): args are AA AB AC AD AF BA BB BC BD BF. */
typedef struct hsm_gate_spec {
    int32_t power;      /* 0 = gate absent; 1..6 */
    int32_t instant;    /* 1 = instantaneous gate (INSTANTX/Y/Z) */
    int32_t mode;       /* 0 = SETUP_ALPHA (B := alpha+beta), 1 = SETUP_TAU */
    int32_t pad;
    double  args[10];
} hsm_gate_spec;

/* A tabchannel prototype.  Copies of one prototype share its tables, as
 * `copy /library/X` does in GENESIS; the hsolve keys tables by pointer
 * (hines_chip_init.c:165-191). */
typedef struct hsm_chantype {
    double        Ek;
    hsm_gate_spec gate[3];  /* X, Y, Z */
} hsm_chantype;

/* Synaptic channel prototype (newconn/synchan.c:535-559 constants). */
typedef struct hsm_syntype {
    double Ek;
    double tau1, tau2;
} hsm_syntype;

typedef struct hsm_spec {
    double  dt;
    int32_t method;         /* HSM_BEULER or HSM_CRANK */
    int32_t chanmode;       /* 4 or 5 */
    int32_t xdivs;          /* gate-table divisions (setupalpha default 3000) */
    int32_t float_fields;   /* 1: element fields are C floats, as GENESIS's
                               compartment/tabchannel structs declare them
                               (segment/seg_struct.h:209-221,
                               sim/struct_defs.h:118-122, olf/olf_struct.h:142-155) */
    double  xmin, xmax;     /* gate-table range (default -0.1 .. 0.05 V) */

    /* compartments in element-list order */
    int32_t ncomp;
    int32_t pad1;
    const int32_t *parent;  /* element index of the axial parent, -1 = root */
    const double  *Cm, *Rm, *Ra, *Em, *initVm, *inject;

    /* tabchannels, in creation order */
    int32_t nchantypes;
    int32_t nchan;
    const hsm_chantype *chantypes;
    const int32_t *chan_comp;   /* element index of the owning compartment */
    const int32_t *chan_type;
    const double  *chan_gbar;

    /* spikegens */
    int32_t nspike;
    int32_t pad2;
    const int32_t *spike_comp;
    const double  *spike_thresh, *spike_refract;

    /* synchans (one SYN2_OP each), in creation order */
    int32_t nsyntypes;
    int32_t nsyn;
    const hsm_syntype *syntypes;
    const int32_t *syn_comp;
    const int32_t *syn_type;
    const double  *syn_gmax;

    /* 1: the compartments are symcompartments (hines_init.c:101-113): the
       elimination program couples siblings through their junction */
    int32_t symmetric;
    int32_t pad3;
} hsm_spec;

/* Emitted hsolve arrays.  Row index = Hines number. */
typedef struct hsm_out {
    int32_t ncompts, nchips, nops, nfuncs, nravals, ncols, xdivs, ngivals;
    int32_t nstab, method, chanmode, nchildren;
    double  xmin, invdx, dt;

    int32_t *ops;        /* [nops]        channel program, LCOMPT_OP-terminated */
    double  *chip;       /* [nchips]      per-compartment chip blocks */
    double  *tablist;    /* [(xdivs+2)*ncols] composite table; row xdivs+1 is a
                            guard copy of row xdivs (see below) */
    double  *xvals;      /* [xdivs+2]     breakpoints; [xdivs+1] is a guard */
    int32_t *funcs;      /* [nfuncs]      elimination program */
    double  *ravals;     /* [nravals]     off-diagonal coefficients */
    double  *vm;         /* [ncompts]     initial voltages (RESET) */
    double  *givals;     /* [ngivals]     chanmode-4 outputs; Rm pre-filled */
    double  *stablist;   /* [4*nstab]     synapse constants */
    int32_t *compops;    /* [ncompts]     ops index of each compartment */
    int32_t *compchips;  /* [ncompts]     chip index of each compartment */
    int32_t *elmnum;     /* [ncompts]     row -> element index */
    int32_t *hnum;       /* [ncompts]     element index -> row */
    int32_t *parent_row; /* [ncompts]     row -> parent row, -1 = root */
    int32_t *spike_row;  /* [nspike]      row of each spikegen (spec order) */
    int32_t *spike_op;   /* [nspike]      ops index of its SPIKE_OP */
    int32_t *syn_op;     /* [nsyn]        ops index of each synchan's SYN2_OP */
    int32_t *syn_chip;   /* [nsyn]        chip index of its X slot */
} hsm_out;

/* Build the hsolve arrays for a spec.  Returns 0, or -1 with a message in
 * hsm_last_error().  Free with hsm_free. */
int32_t hsm_build(const hsm_spec *spec, hsm_out **out);
void    hsm_free(hsm_out *out);
const char *hsm_last_error(void);

/* Interpolated lookup of an untransformed gate table, TabInterp LIN_INTERP
 * (olf/new_interp.c:208-272).  Exposed for tests. */
double  hsm_tabinterp(const double *table, int32_t xdivs, double xmin,
                      double invdx, double x);

/* ---- synthetic populations ------------------------------------------- */

/* Random Hines trees with HH soma and Kdr dendrites (the synthetic
 * workloads).  sizes[] / probs[] give the compartment-count distribution. */
typedef struct hsm_pop {
    int32_t n_neurons;
    int32_t nsizes;
    const int32_t *sizes;
    const double  *probs;    /* cumulative sampling uses these weights */
    uint64_t seed;
    double dt;
    int32_t method;
    int32_t max_kids;        /* parent resampled while it has this many kids */
    double soma_len, soma_dia;
    double dend_len_lo, dend_len_hi, dend_dia_lo, dend_dia_hi;
    double CM, RM, RA, Em;   /* specific membrane parameters */
    double gNa, gK, gKdr;    /* S/m^2 */
    double inject_lo, inject_hi;
    double erest_act;        /* HH table rest potential */
    int32_t float_fields;    /* as hsm_spec.float_fields */
    int32_t first_neuron;    /* global index of the first neuron built */
} hsm_pop;

int32_t hsm_build_population(const hsm_pop *pop, hsm_out **out);

#ifdef __cplusplus
}
#endif

#endif
