/*
 * hsolve_hip.h -- C ABI of libhsolve_hip, the MI355X (gfx950) backend for the
 * chanmode-4/5 step of GENESIS's hsolve Hines solver.
 *
 * The library takes the arrays an hsolve holds after SETUP/RESET (the ops[]
 * channel program, chip[] values, the composite gate table, the funcs[] /
 * ravals[] elimination program and the Hines-ordered vm[]), copies them once
 * into device-resident structure-of-arrays, and advances the solver on the
 * GPU.  Plain pointers and sizes only; no GENESIS or torch types.
 *
 * Reference interfaces replaced (genesis/src/hines/cuda/):
 *   hsh_create       cuda_backend_init (cuda_backend.cu:157-259) +
 *                    cuda_backend_tree_init (cuda_backend.cu:396-444), and the
 *                    refusal scan of cuda_init (cuda_hsolve.c:289-379)
 *   hsh_refusals     the stderr opcode report of cuda_init (cuda_hsolve.c:321-341)
 *   hsh_step         cuda_backend_perstep (cuda_backend.cu:263-316),
 *                    cuda_backend_multiloop (:338-385) and
 *                    cuda_backend_multiloop_tree (:461-527); return protocol of
 *                    cuda_chip_update (cuda_hsolve.c:381-495, hines.c:225-252)
 *   hsh_get_state    cuda_backend_sync_chip (cuda_backend.cu:529-538) /
 *                    cuda_sync_chip (cuda_hsolve.h:24)
 *   hsh_put_state    cuda_backend_upload_chip (cuda_backend.cu:318-336)
 *   hsh_set_inputs   h_in_msgs (hines_msgs.c:351-405): message inputs before a step
 *   hsh_set_chip     its chip overwrites alone
 *   hsh_destroy      cuda_backend_cleanup (cuda_backend.cu:540-568)
 *   hsh_autotune     no counterpart: the reference multiloop launches one fixed
 *                    shape (ceil(n_trees/64) x 64, cuda_backend.cu:461-527)
 *
 * Conventions: statuses are hsh_status codes (negative = error; HSH_REFUSED is
 * not an error: the caller falls back to its CPU path for this solver).  No
 * function calls exit().  One handle per hsolve; a handle is not thread-safe,
 * distinct handles are.  Each handle owns one HIP stream and its graphs.
 */
#ifndef HSOLVE_HIP_H
#define HSOLVE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hsh_status {
    HSH_OK = 0,
    HSH_REFUSED = 1,         /* model holds something the device path cannot run */
    HSH_ERR_ARG = -1,        /* bad argument or malformed arrays */
    HSH_ERR_DEVICE = -2,     /* HIP runtime error (message in hsh_last_error) */
    HSH_ERR_NOMEM = -3,
    HSH_ERR_STATE = -4,      /* call not valid in this state (e.g. refused handle) */
    HSH_ERR_LIMIT = -5       /* a size exceeds a documented maximum */
} hsh_status;

/* integration methods (genesis/src/sim/sim_defs.h:174-175) */
#define HSH_BEULER 10
#define HSH_CRANK  11

/* Hsolve-format input.  All arrays are host pointers, copied at hsh_create;
 * the caller keeps ownership.  Sizes and layouts are those of struct
 * hsolve_type (genesis/src/hines/hines_struct.h:126-417). */
typedef struct hsh_model {
    int32_t ncompts;         /* rows, Hines order */
    int32_t nchips;
    int32_t nops;            /* ops[] length including the LCOMPT_OP terminator */
    int32_t nfuncs;
    int32_t nravals;
    int32_t ncols;           /* composite gate-table columns */
    int32_t xdivs;           /* table rows are xdivs+1 (plus one guard row) */
    int32_t nstab;           /* synapse types; stablist holds 4*nstab values */
    int32_t method;          /* HSH_BEULER or HSH_CRANK */
    int32_t chanmode;        /* 4 or 5 */
    int32_t calcmode;        /* 1 = interpolated tables (the only supported mode) */
    int32_t nconcs;          /* entries of conc[] (0: no concentrations) */
    double  dt;
    double  xmin;
    double  invdx;
    const int32_t *ops;      /* [nops] */
    const double  *chip;     /* [nchips] */
    const double  *tablist;  /* [(xdivs+2)*ncols]: row xdivs+1 is a guard row that
                                repeats row xdivs (the reference reads that row
                                when Vm >= xmax, with interpolation weight 0) */
    const double  *xvals;    /* [xdivs+2] */
    const double  *stablist; /* [4*nstab] or NULL */
    const int32_t *funcs;    /* [nfuncs] */
    const double  *ravals;   /* [nravals] */
    const double  *vm;       /* [ncompts] */
    /* Concentrations (hines_struct.h:315-376); all zero / NULL without
     * them.  The concentration-indexed rows of the gate table share tablist
     * with the voltage rows: with ydivs > 0 tablist holds
     * (max(xdivs, ydivs)+2)*ncols values, the last row a guard row that
     * repeats row max(xdivs, ydivs). */
    int32_t ydivs;           /* conc table rows are ydivs+1; a row is found from
                                ymin/invdy (hines_chip.c:238-247), so yvals[] is
                                not needed */
    int32_t nconcops;        /* concops[] length including its LCONC_OP / LCACONC_OP */
    int32_t nconcchip;
    double  ymin;
    double  invdy;
    const double  *conc;     /* [nconcs] concentrations: pools, shells, buffers */
    const int32_t *concops;  /* [nconcops]: the concentration program of
                                h_init_conc_chip (hines_conc_init.c:413-805) --
                                CONC_OP / FCONC_OP k, FLUXCONC_OP, TAUPUMP_OP row,
                                CTAUPUMP_OP, MMPUMP_OP, HILLPUMP_OP hill,
                                PHILLPUMP_OP hill shell, BUFFER_OP buffer, POOL_OP,
                                CACONC_OP k -- ended by LCONC_OP or LCACONC_OP */
    const double  *concchip; /* [nconcchip]: the values those ops consume, in order
                                (hines_data.c:162-400) */
    /* chanmode-4/5 output stream (hines_chip.c:164-165, :536-564, :598): when
     * given, the device writes it every step -- per compartment the leak
     * current, {Gk, Ik, Ek} of every current (chanmode 5: Gk and Ik per gmax)
     * and the total membrane current; the Rm entries the reference reads are
     * taken from here.  NULL / 0: not produced (no cost). */
    int32_t ngivals;
    const double  *givals;   /* [ngivals] */
    /* Diffusion systems of shells and buffers, the solve of do_h_conc_solve
     * (hines_conc.c:39-396; arrays of h_init_conc_solve / h_setup_conc_solve,
     * hines_conc_init.c:59-411).  System i: nshells[i] shells (at least two)
     * of nshrows[i] rows -- its nshfix[i] fixed buffers, its diffusible
     * buffers, the shell last -- with row r of shell m at conc[dconc[i] +
     * m*nshrows[i] + r] (shell 0 innermost), buffer rates at
     * bufvals[fbufval[i]] (three per buffer).  covals0 holds every system's
     * couplings and fixed-buffer totals in the order the solve consumes them.
     * fdiffs[ncompts+1] (hines_struct.h fdiffs): the systems of compartment
     * (Hines row) r are fdiffs[r] .. fdiffs[r+1]-1.  ndiffs = 0: none. */
    int32_t ndiffs;
    int32_t ncovals0;
    int32_t nbufvals;
    int32_t pad_diff;
    const int32_t *nshells;  /* [ndiffs] */
    const int32_t *nshrows;  /* [ndiffs] */
    const int32_t *nshfix;   /* [ndiffs] */
    const int32_t *dconc;    /* [ndiffs] */
    const int32_t *fbufval;  /* [ndiffs] */
    const int32_t *fdiffs;   /* [ncompts+1] */
    const double  *covals0;  /* [ncovals0] */
    const double  *bufvals;  /* [nbufvals] */
} hsh_model;

typedef struct hsh_options {
    int32_t device;          /* HIP device ordinal */
    int32_t batch_k;         /* steps per batched launch; 0 = autotuned */
    int32_t batch_waves;     /* waves per workgroup of the tile kernel for trees of
                                17..64 rows; 0 = rows/4 */
    int32_t flags;           /* HSH_F_* */
    /* Layout.  Trees of up to tile_max_rows rows run in 64-neuron tiles (one
     * lane per neuron); larger trees are packed, whole, into workgroup-sized
     * packs (one lane per compartment, level-scheduled elimination). */
    int32_t tile_max_rows;   /* 0 = autotuned; < 0 = every tree in packs */
    int32_t pack_cap;        /* compartments per pack; 0 = autotuned */
    int32_t pack_waves;      /* waves per pack workgroup (1..8); 0 = autotuned */
    int32_t pack_stage;      /* pack schedule in LDS: 0 = autotuned, 1 = no, 2 = all of it, 3 = its indices */
} hsh_options;

#define HSH_F_NO_AUTOTUNE 1  /* no autotuning: options as given, defaults for the rest
                                (K 32, tile_max_rows 16, pack_cap 1024, 4 pack waves,
                                schedule not staged) */
#define HSH_F_HOST_ONLY   2  /* analyse and plan only, allocate nothing on a device:
                                hsh_refusals and hsh_info_get work, stepping and
                                state calls return HSH_ERR_STATE */
#define HSH_F_NO_JIT      4  /* use the built-in micro-op interpreter kernels instead
                                of kernels specialised to the model's channel programs
                                and compiled at hsh_create (hipRTC) */

/* refusal reasons */
#define HSH_R_OPCODE         1   /* opcode the device interpreter does not run */
#define HSH_R_EK_LIVEIN      2   /* channel inherits a computed Ek from an earlier compartment */
#define HSH_R_TABLE_LIVEIN   3   /* gate lookup without NEWVOLT in its compartment */
#define HSH_R_REGISTER_LIVEIN 4  /* other register carried in from an earlier compartment */
#define HSH_R_FUNCS          5   /* elimination program the device solver does not run: an
                                    operand out of range, or a row cursor that leaves
                                    the solver (sibling programs of symcompartments run
                                    as units after the host has walked them) */
#define HSH_R_TREE           6   /* tree rows are not contiguous / kid order not ascending */
#define HSH_R_CALCMODE       7   /* calcmode 0 (no interpolation) */
#define HSH_R_SYN_PENDING    8   /* SYN2_OP counter holds a pending event (>= 0) */
#define HSH_R_POOL           9   /* concentration pool fed or read by several compartments,
                                    updated with no compartment feeding it, or more
                                    than 4 pools in one compartment */
#define HSH_R_DIFFUSION     10   /* a diffusion system or concentration op the device does
                                    not run: fewer than two shells, arrays out of range,
                                    a concentration program whose owner compartment
                                    cannot be told (a concpool not fed by a Hill pump
                                    from a shell), a TAUPUMP_OP reading another
                                    compartment's Vm (a group with such solvers is
                                    HSH_ERR_ARG at hsh_create_group) */

typedef struct hsh_refusal {
    int32_t comp;            /* row (Hines number), -1 for whole-solver reasons */
    int32_t opcode;          /* offending ops[]/funcs[] code, 0 if none */
    int32_t op_index;        /* index into ops[] or funcs[], -1 if none */
    int32_t reason;          /* HSH_R_* */
} hsh_refusal;

/* step modes */
#define HSH_PLAIN    0       /* per-step drop-in: channel kernel then elimination
                                kernel, one step per pair of launches */
#define HSH_BATCHED  1       /* persistent fused kernel, K steps per launch, replayed
                                from a captured hipGraph */
#define HSH_CHANNELS 2       /* channel interpreter only: results[] on the device,
                                vm untouched (reference return-0 protocol).  A
                                compartment's concentration updates (Ca_concen
                                pools, its concentration ops and diffusion systems)
                                run inside its program, so they advance here too:
                                a caller keeping the CPU elimination skips its own
                                do_h4_conc_chip_update / do_h_conc_solve */

typedef struct hsh_step_out {
    int32_t vm_ready;        /* 1: vm[] is final for these steps (skip the CPU
                                elimination); 0: only results[] were produced */
    int32_t nspikes;         /* spike events raised during these steps */
    int32_t launches;        /* kernel launches (or graph replays) issued */
    int32_t pad;
} hsh_step_out;

typedef struct hsh_info {
    int32_t ntiles;          /* 64-lane tiles of neurons */
    int32_t nblocks;         /* (tile,row) blocks */
    int32_t ntemplates;      /* distinct channel programs */
    int32_t ntrees;          /* neurons */
    int32_t max_rows;        /* largest neuron in a tile */
    int32_t batch_k;         /* chosen K */
    int32_t batch_waves;     /* chosen waves per workgroup */
    int32_t refused;         /* 1 if the handle is refused */
    double  bytes_per_step;  /* algorithmic HBM bytes of one plain step */
    double  lane_fill;       /* real compartments / (blocks*64) */
    int32_t jit;             /* 1: specialised kernels in use */
    int32_t npacks;          /* packs of larger trees */
    double  jit_compile_s;   /* time spent generating + compiling them */
    int32_t tile_max_rows;   /* layout in use (see hsh_options) */
    int32_t pack_cap;
    int32_t pack_waves;
    int32_t pack_stage;      /* 1 = no, 2 = all of it, 3 = its indices */
    int32_t max_tree_rows;   /* largest neuron */
    int32_t autotuned;       /* 1 once the autotuner has run */
    double  autotune_s;      /* time it took */
} hsh_info;

/* opaque solver handle */
struct hsh_handle;

/* Where one solver of a group sits in the group's arrays (hsh_get_state and
 * friends return the members' arrays concatenated in member order). */
typedef struct hsh_member {
    int32_t row0, nrows;     /* rows in vm[] (and results[], two per row) */
    int32_t chip0, nchips;   /* chip[] */
    int32_t op0, nops;       /* ops[], without the member's LCOMPT_OP */
    int32_t conc0, nconcs;   /* conc[] */
    int32_t giv0, ngivals;   /* givals[] */
    int32_t stab0;           /* its stablist entries start here in the group's (the
                                stab of an hsh_input for one of its synchans is its
                                own index plus stab0) */
    int32_t pad;
} hsh_member;

hsh_status hsh_create(const hsh_model *m, const hsh_options *opt, struct hsh_handle **out);
/* Several solvers (e.g. one hsolve per cell, the reference's VAnet2 layout)
 * advanced by one handle: one launch per step for all of them, and synaptic
 * connections between them (hsh_set_synapses with group ops indices).  Each
 * member keeps the semantics of its own solver -- interpreter registers and
 * row caches start afresh at its first compartment, its funcs[]/ravals[] are
 * its own -- so vm, chip, ops, conc and givals equal the members' separate
 * runs.  results[] follow the one-solver convention: the last row and
 * single-compartment solvers of all but the last member hold their solved
 * values.  Members must share dt, method, chanmode, calcmode, table axes
 * (xdivs/xmin/invdx, ydivs/ymin/invdy) and whether they carry givals; equal
 * tables and synapse tables are stored once.  Refusal rows are group rows.
 * No reference counterpart: the reference batches one-cell solvers on the
 * host each step (cuda_hsolve.c:65-145) and not with spikegens. */
hsh_status hsh_create_group(const hsh_model *const *models, int32_t n, const hsh_options *opt,
                            struct hsh_handle **out);
hsh_status hsh_group_member(struct hsh_handle *h, int32_t i, hsh_member *out);
hsh_status hsh_refusals(struct hsh_handle *h, hsh_refusal *out, int32_t cap, int32_t *n);
hsh_status hsh_step(struct hsh_handle *h, int32_t nsteps, int32_t mode, hsh_step_out *out);
hsh_status hsh_sync(struct hsh_handle *h);
/* Any pointer may be NULL.  Arrays are in the reference layout:
 * vm[ncompts], chip[nchips], results[2*ncompts], ops[nops]. */
hsh_status hsh_get_state(struct hsh_handle *h, double *vm, double *chip, double *results,
                         int32_t *ops);
hsh_status hsh_put_state(struct hsh_handle *h, const double *vm, const double *chip,
                         const int32_t *ops);
/* Pool concentrations conc[nconcs] (reference indices).  The Ca_concen
 * update runs on the device in the program of the compartment that feeds the
 * pool (do_h4_conc_chip_update, hines.c:239-240, hines_conc_chip.c:196-207). */
hsh_status hsh_get_conc(struct hsh_handle *h, double *conc);
hsh_status hsh_put_conc(struct hsh_handle *h, const double *conc);
/* givals[ngivals] as the last step wrote it (HSH_ERR_STATE if the model was
 * created without givals). */
hsh_status hsh_get_givals(struct hsh_handle *h, double *givals);
/* Message inputs, the h_in_msgs step of PROCESS (hines_msgs.c:351-405),
 * applied to the device state before the next step, in the order given (the
 * reference's ininfo walk).  op and index as the reference's ininfo entry:
 *   HSH_IN_INJECT / HSH_IN_BASE / HSH_IN_THRESH: chip[index] = value;
 *   HSH_IN_ACTIVATION: an ACTIVATION message to the synchan whose X is
 *     chip[index], of synapse type stab (the stablist index its SYN2_OP
 *     holds): if value > VTINY (1e-10), X += value*stablist[stab+3], times
 *     dt when per_dt (a synchan's ACTIVATION, not a channelC2/C3's);
 *   HSH_IN_FLUX: flux[index] = value -- accepted and without effect: in
 *     chanmodes 4 and 5 CONC_VAL_OP clears the pool's flux at the start of
 *     every program that feeds it (hines_chip.c:205-212), and the device
 *     runs only pools such a program feeds.  A FLUX input into a shell
 *     whose FLUXCONC_OP reads a flux no CONC_VAL_OP clears is HSH_ERR_ARG
 *     (the device takes that flux as its initial 0).
 * Indices are those of the handle's arrays (a group's merged arrays). */
#define HSH_IN_INJECT     1     /* INJECT_OP (hines_defs.h:225) */
#define HSH_IN_BASE       2     /* BASE_OP */
#define HSH_IN_THRESH     3     /* THRESH_OP */
#define HSH_IN_FLUX       3300  /* FLUX_OP */
#define HSH_IN_ACTIVATION 4100  /* SYN2_OP */
typedef struct hsh_input {
    int32_t op;
    int32_t index;
    int32_t stab;
    int32_t per_dt;
    double  value;
} hsh_input;
hsh_status hsh_set_inputs(struct hsh_handle *h, int32_t n, const hsh_input *in);
/* hsh_set_inputs that allocates nothing: its device staging is a
 * fixed-size array of static storage in the library (one per GPU, kept
 * until the process ends) and its host staging belongs to the handle.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (serving loops, harnesses that forbid it).  At most
 * HSH_INPUTS_STATIC_MAX entries after expansion (an overwrite counts once
 * per device slot holding the chip value); more is HSH_ERR_LIMIT with
 * nothing applied.  Calls that overlap on one GPU (other threads, other
 * handles on that GPU) share the array and race; a call returns only after
 * its inputs are applied. */
#define HSH_INPUTS_STATIC_MAX 16384
hsh_status hsh_set_inputs_static(struct hsh_handle *h, int32_t n, const hsh_input *in);
/* Overwrite n chip[] values (reference indices), e.g. INJECT messages
 * (hsh_set_inputs with HSH_IN_INJECT). */
hsh_status hsh_set_chip(struct hsh_handle *h, int32_t n, const int32_t *idx, const double *val);
/* Spike events raised by SPIKE_OP since the last drain: ops[] index of the
 * SPIKE_OP and the step number. */
hsh_status hsh_drain_spikes(struct hsh_handle *h, int32_t *opidx, int32_t *step, int32_t cap,
                            int32_t *n);
hsh_status hsh_info_get(struct hsh_handle *h, hsh_info *info);
/* Run the autotuner now (otherwise it runs before the first HSH_BATCHED step
 * unless HSH_F_NO_AUTOTUNE): times candidate layouts, pack kernel shapes and
 * K on the handle's own model and keeps the fastest.  The simulation state
 * (vm, chip, ops, results, step count, pending spikes) is unchanged. */
hsh_status hsh_autotune(struct hsh_handle *h);
/* Synaptic connections and the synchans' pending event lists, on the device.
 * Each connection takes a spike raised by the SPIKE_OP at ops index
 * src_op[i] to the SYN2_OP synchan at ops index dst_op[i] with weight[i] and
 * delay[i] (s), rounded to float as GENESIS stores them
 * (newconn_struct.h:158-159).  Every synchan keeps the reference's pending
 * list (newconn/synchan.c:161-358): a spike is an EVENT that hits
 * (int)(delay/dt + 0.9) steps later, timed relative to the event before it;
 * a non-empty list first re-times its head by (int)((t - time_last_event)/dt)
 * of the floating simulation times; a new head sets the SYN2_OP counter
 * (HSEVENT, hines.c:352-357); at 0 the synchan sums weight/dt of the events
 * of that step into its float activation and adds dt*activation*
 * stablist[k+3] to X (h_dosynchan, hines_conn.c:60-89).  EVENTs within a
 * step follow ops[] order (a synchan whose op comes before the spike's has
 * already counted that step).  Replaces the reference's per-spike host
 * messaging (h_dospike_event, hines_conn.c:44-56).  Batched launches run at
 * most (shortest connection in steps) - 1 steps, and at most 64.  n = 0
 * removes the connections; pending events stay.  HSH_ERR_LIMIT later (from
 * hsh_sync / hsh_get_state) if a list overflows its capacity (the events one
 * connection can hold pending over its delay, plus 32). */
hsh_status hsh_set_synapses(struct hsh_handle *h, int32_t n, const int32_t *src_op,
                            const int32_t *dst_op, const double *weight, const double *delay);
/* EVENTs that reach hsolved synchans from outside the handle: spikegens and
 * other solvers connected through GENESIS messages (hines_conn.c:91-105 ->
 * synchan EVENT -> HSEVENT, hines.c:352-357).  steps[i] is the caller's
 * (int)((delay - (SimulationTime() - eventtime))/dt + 0.9); weight[i] is the
 * synapse weight (rounded to float).  after = 0: the events arrive at the
 * start of the next step, before the solver's ops (their source ran earlier
 * in the step); after = 1: at the end of the last step run, after its ops.
 * Within one call, and across calls before the next step, events keep the
 * order they were posted in.  They are applied by the next hsh_step (or
 * hsh_get_state).  Replaces the reference's synchan EVENT path for hsolved
 * synchans; no counterpart in the reference's GPU backend (which has none). */
hsh_status hsh_post_events(struct hsh_handle *h, int32_t n, const int32_t *syn_op,
                           const int32_t *steps, const double *weight, int32_t after);
/* SimulationTime() of the next step (0 at hsh_create); it then grows by dt
 * per step as simulation_time += ClockValue(0) does (sim/sim_step.c:112).
 * The pending lists' re-timing reads it. */
hsh_status hsh_set_time(struct hsh_handle *h, double t);
/* Replace the elimination coefficients with ravals[nravals] in the layout of
 * the model's funcs[] program (the values h_funcs_init computes from Cm and
 * Ra, genesis/src/hines/hines_init.c:604-742, :1011-1172); the state is kept. */
hsh_status hsh_set_ravals(struct hsh_handle *h, const double *ravals);
/* Adjoint (reverse-mode gradients) of a run, for parameter fitting.  No
 * counterpart in the reference.
 * hsh_adj_forward runs nsteps plain steps from the current state, keeps a
 * checkpoint of the state every ckpt steps, and writes vm of the rows obs[]
 * (distinct) after every step to trace[nsteps*nobs] (host, may be NULL).  The
 * state advances as with hsh_step.  Tabchannel programs only (up to three
 * gates per channel); HSH_REFUSED otherwise.
 * hsh_adj_backward takes dL/dtrace[nsteps*nobs] for the last forward run and
 * writes dL/dchip[nchips], dL/dravals[nravals] and dL/dvm0[ncompts] (any may
 * be NULL): constants of the channel programs (gmax, Ek, Em/Rm, inject, tbyc,
 * diagterm) and the coefficients summed over the run, gate states and vm at
 * its start.  It recomputes each segment from its checkpoint (memory: one
 * state per checkpoint plus ckpt states) and leaves the state as the forward
 * run left it. */
hsh_status hsh_adj_forward(struct hsh_handle *h, int32_t nsteps, int32_t ckpt, int32_t nobs,
                           const int32_t *obs, double *trace);
hsh_status hsh_adj_backward(struct hsh_handle *h, const double *dtrace, double *g_chip,
                            double *g_ravals, double *g_vm0);
/* Device time of the last hsh_step call (HIP events on the handle's stream),
 * in ms, after hsh_sync.  kernel selects HSH_T_*. */
#define HSH_T_TOTAL    0
#define HSH_T_CHANNEL  1     /* plain mode: channel kernel only */
#define HSH_T_ELIM     2     /* plain mode: elimination kernel only */
hsh_status hsh_last_time(struct hsh_handle *h, int32_t kernel, double *ms);
/* Enable per-kernel event timing for following hsh_step calls (adds events
 * around every launch). */
hsh_status hsh_set_timing(struct hsh_handle *h, int32_t on);
/* Diagnostic: the device source libhsolve_hip generates for this model's
 * channel programs (what hsh_create compiles with hipRTC).  Writes up to cap
 * bytes (NUL-terminated) and the full length to *len.  Needs no device. */
hsh_status hsh_jit_source(const hsh_model *m, char *buf, int64_t cap, int64_t *len);
void       hsh_destroy(struct hsh_handle *h);
const char *hsh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
