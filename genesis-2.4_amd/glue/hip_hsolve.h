/*
 * hip_hsolve.h -- GENESIS-side entry points of the MI355X hsolve backend.
 *
 * Same role and return protocol as the CUDA backend's header
 * (genesis/src/hines/cuda/cuda_hsolve.h:22-25): hines.c selects it with
 *     #define ACCEL_CHIP_UPDATE hip_chip_update
 * under USE_HIP (genesis/src/hines/hines.c:93-98).
 */
#ifndef HIP_HSOLVE_H
#define HIP_HSOLVE_H

#ifdef USE_HIP
typedef struct hsolve_type Hsolve;

int  hip_chip_update(Hsolve *hsolve);   /* 1: vm[] final (skip the CPU solve), 0: CPU path ran */
void hip_sync_chip(Hsolve *hsolve);     /* device state -> vm/chip/results/ops (HGET, SAVE2) */
void hip_release(Hsolve *hsolve);       /* free this solver's device state (h_delete) */
int  hip_owns_conc(Hsolve *hsolve);     /* 1: the device updates conc[] (skip do_h4_conc_chip_update) */
void hip_push_state(Hsolve *hsolve);    /* host vm/chip/ops/conc -> device (HPUT, HCHANGE, HRESTORE, RESTORE2) */
/* From the synchan EVENT action (newconn/synchan.c:193-194) of an hsolved
 * synchan: 1 if its solver runs on the device and took the event (the
 * element's own pending list is then left alone), 0 for the element path. */
int  hip_synchan_event(Element *hsolve, int solve_index, int time, double weight);
#endif

#endif
