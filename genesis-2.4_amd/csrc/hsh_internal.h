// hsh_internal.h -- shared definitions of libhsolve_hip (host + device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "hsolve_hip.h"
#include "hsh_devtypes.h"

namespace hsh {

// Reference opcodes (genesis/src/hines/hines_defs.h:178-222).
enum RefOp : int32_t {
    COMPT_OP = 100, FCOMPT_OP = 101, LCOMPT_OP = 102, SPIKE_OP = 200, ELMMPUMP_OP = 500,
    CONC_VAL_OP = 1000, NERNST_OP = 2000, GHK_OP = 2010, CHAN_OP = 3000, CHAN_EK_OP = 3001,
    SAME_GHK_OP = 3002, DO_TABGHK_OP = 3003, ADD_CURR_OP = 3200, FLUX_OP = 3300,
    SCALE_FLUX_OP = 3301, STORE_ALLI_OP = 3501, STORE_ALLC_OP = 3502, IPOL1V_OP = 4001,
    IPOL1C_OP = 4002, IPOL2V_OP = 4003, IPOL2C_OP = 4004, SYN2_OP = 4100, SYN3_OP = 4101,
    MGBLOCK_OP = 4200, NEWVOLT_OP = 5100, NEWCONC1_OP = 5110, NEWCONC2_OP = 5111,
    DOMAIN_OP = 5112, CONC_OP = 10100, FCONC_OP = 10101, LCONC_OP = 10102, FLUXCONC_OP = 10103,
    CACONC_OP = 10200, LCACONC_OP = 10202, TAUPUMP_OP = 11000, CTAUPUMP_OP = 11001, MMPUMP_OP = 11010,
    HILLPUMP_OP = 11011, PHILLPUMP_OP = 11012, BUFFER_OP = 12000, POOL_OP = 12100,
};
// Reference funcs codes (hines_defs.h:119-128).
enum RefFunc : int32_t {
    F_ELIM = 0, F_SET_DIAG = 1, F_SKIP_DIAG = 2, F_FASTSIB = 3, F_COPY = 4, F_SIB = 5,
    F_CALC = 6, F_FINISH = 7,
};

struct Template {
    std::vector<int32_t> code;
    int nslots = 0;
    int nislots = 0;
    int ngates = 0;
    int nscratch = 0;     // slots that only carry values within a step (the
                          // diffusion solve's diagonals, results and elimination)
};

// Layout choices (the autotuner sweeps them).
struct LayoutOpts {
    int tile_max_rows = 16;   // larger trees go to packs (0: every tree)
    int pack_cap = 1024;      // compartments per pack (a larger tree fills one alone)
};

// Host-side result of analysing a model.
struct Plan {
    std::vector<hsh_refusal> refusals;
    bool giv = false;                     // produce givals (chanmode-4/5 output stream)
    bool chanmode5 = false;
    std::vector<Template> templates;
    // per compartment (Hines row)
    std::vector<int32_t> comp_tmpl;
    std::vector<int64_t> comp_dsrc_off;   // offset into dsrc
    std::vector<int64_t> comp_isrc_off;
    // device slot sources: >=0 reference chip index; <0: -(1+k) synthetic value k
    std::vector<int64_t> dsrc;
    std::vector<double> synth;
    // int slot sources: >=0 reference ops index (mutable/const ints); <0: -(1+v) literal v
    std::vector<int64_t> isrc;
    // tree
    std::vector<int32_t> parent;          // parent row (global), -1 root
    std::vector<double> ta, tcf, tcb;
    std::vector<int32_t> ra_fwd;          // ravals index of (ta, tcf) of a non-root row, -1
    std::vector<int32_t> ra_bwd;          // ravals index of tcb of a non-root row, -1
    std::vector<int32_t> tree_lo;         // per tree (neuron): first row; root = hi
    std::vector<int32_t> tree_hi;
    std::vector<int32_t> pred;            // previous NEWVOLT compartment (global row), -1
    std::vector<int32_t> spike_ops;       // spike source id -> ops index of its SPIKE_OP
    std::vector<int32_t> spike_comp;      // spike source id -> its compartment (row)
    std::vector<int64_t> spike_thresh;    // spike source id -> chip index of its threshold
    std::vector<int32_t> syn_ops;         // synchan id -> ops index of its SYN2_OP
    std::vector<int32_t> prog;            // programs of all templates
    std::vector<int32_t> tprog;           // template -> offset in prog
    int max_tree_rows = 0;
    double bytes_per_step = 0;
    // per concentration: a FLUXCONC_OP reads its flux with no CONC_VAL_OP of
    // the owner's program clearing it first (the device takes that flux as
    // its initial 0; a FLUX message input would change it)
    std::vector<char> flux_live;
    // layout (build_layout)
    LayoutOpts layout;
    std::vector<BlockMeta> blocks;
    std::vector<TileMeta> tiles;
    std::vector<PackMeta> packs;
    std::vector<int32_t> lvl, fent, bkp;
    std::vector<uint16_t> kidx;
    std::vector<double> kta, ktcf, btcb;
    std::vector<int32_t> comp_pos;        // row -> block*64+lane
    std::vector<int32_t> pos_comp;        // block*64+lane -> row (padding lanes: lane-0 row)
    int64_t nS = 0, nI = 0;
    int max_rows = 0;                     // rows of the largest tiled tree
    int max_pack_pos = 0;                 // positions of the largest pack
    // Solvers of symmetric compartments: their elimination program couples
    // siblings (COPY_ARRAY, SIBARRAY_ELIM, FASTSIBARRAY_ELIM), so each runs
    // as one unit, one lane interpreting its funcs[] over a mutable ravals
    // copy.  Its rows form one tree of the layout (a chain of placeholder
    // parents), placed in tiles after the other trees.
    struct FuncsUnit {
        int32_t r0, nn;                   // rows
        int64_t f0, nf;                   // its program in ffuncs
        int64_t q0, nq;                   // its ravals in fravals
        int64_t qg;                       // and in the caller's (merged) ravals
    };
    std::vector<FuncsUnit> funits;
    std::vector<int32_t> ffuncs;
    std::vector<double> fravals;
    std::vector<int32_t> tree_funit;      // per tree: its unit, -1
    std::vector<int32_t> funit_tile, funit_lane;   // per unit: tile and lane (build_layout)
    int ntiles_regular = 0;               // tiles before the units' tiles
    // LDS bytes of the pack kernels: schedule in memory, all of it in LDS, its indices in LDS
    int64_t pack_lds = 0, pack_lds_stage = 0, pack_lds_stage_idx = 0;
};

// Several solvers in one handle (hsh_create_group): the analysed model is
// their arrays concatenated, and each member keeps its own semantics -- the
// interpreter registers start afresh at its first compartment, its operands
// (table columns, synapse-table entries, pools) are offset into the merged
// arrays, and its funcs[]/ravals[] are decoded on their own.
struct GroupSpec {
    std::vector<int32_t> member_of;                   // merged row -> member
    std::vector<int32_t> row0, col0, stab0, conc0, giv0, ra0;
    std::vector<const int32_t*> funcs;
    std::vector<int32_t> nfuncs, nravals, nrows;
    std::vector<const double*> ravals;
};

// Analyse ops[] / funcs[] and build the device layout.  Returns false on a
// malformed model (message in err); refusals go into plan.refusals.
bool analyse(const hsh_model& m, Plan& plan, std::string& err, const GroupSpec* g = nullptr);
// (Re)build the device layout of an analysed model.
bool build_layout(Plan& plan, const LayoutOpts& o, std::string& err);

// Kernels specialised to a model's channel programs and compiled at run time
// (hsh_jit.cpp).  ok == false: the built-in interpreter kernels are used.
struct JitModule {
    hipModule_t mod = nullptr;
    hipFunction_t channel = nullptr, batched_r[3] = {nullptr, nullptr, nullptr}, batched = nullptr,
                  iso = nullptr, forest[3] = {nullptr, nullptr, nullptr}, adj_forest = nullptr;
    bool ok = false;
    double compile_s = 0.0;
};
std::string jit_source(const Plan& P, int ncols);
bool jit_build(const Plan& P, int ncols, JitModule& jm, std::string& err);
void jit_free(JitModule& jm);

// Synchans' pending event lists on the device (hsh_events.hip).  Each step
// first runs the spike pass: every SPIKE_OP of the model (it reads only Vm,
// its threshold and its counter, so it can run before the step's other
// ops), recording spikes and setting the sources' flags; then the step's
// kernels run, and each SYN2_OP keeps its own synchan's list from the flags
// (syn2_op, hsh_device.h).
constexpr int EV_ERR_FULL = 2;    // a pending list ran out of room

struct EvPost {
    const int* sid;       // sorted by synchan, then post order
    const int* T;
    const double* wdt;
    const int* after;
    const int* ipos;      // per synchan: I position of its SYN2_OP counter
    int nsyn;
};

// The spike pass of the step every counter holds, over d.src (d.nsrc sources).
hipError_t launch_ev_spikes(const DevModel& d, hipStream_t s);
hipError_t launch_ev_post(const DevModel& d, const EvPost& p, EvCtl* ctl, hipStream_t s);

// A message input applied to a device slot before a step (h_in_msgs,
// hines_msgs.c:351-405): mode 0 sets S[pos] = value, mode 1 adds value.
// Entries are sorted by position, in the caller's order within one.
// A unit of hsh_kernels.hip's funcs[] interpreter: tile and lane of its rows
// (row r at block tiles[tile].block0 + r), its program and ravals offsets.
struct FuncsUnitDev {
    int tile, lane, nn, pad;
    long long f0, q0;
};
hipError_t launch_funcs(const DevModel& d, const FuncsUnitDev* u, int nu, const int32_t* ffuncs, double* fwork,
                        hipStream_t s);

struct InEntry {
    int pos;
    int mode;
    double value;
};
hipError_t launch_inputs(double* S, const InEntry* e, int n, hipStream_t s);
// The static staging array of hsh_set_inputs_static: look up its address on
// a device (with that device current; once per device), and get it.
hipError_t inputs_static_setup(int device);
InEntry* inputs_static_ptr(int device);

// kernel launchers (hsh_kernels.hip); jm may be null or not ok
hipError_t launch_channel(const DevModel& d, const JitModule* jm, hipStream_t s);
hipError_t launch_elim(const DevModel& d, int max_rows, hipStream_t s);
// waves == 0 selects the four-wave register-coefficient kernel (rows <= 16)
hipError_t launch_batched(const DevModel& d, const JitModule* jm, int ksteps, int waves, int max_rows,
                          hipStream_t s);
hipError_t launch_iso(const DevModel& d, const JitModule* jm, int ksteps, hipStream_t s);
// packs: batched (K steps, channel programs + elimination), waves per pack,
// stage = schedule copied to LDS (0 none, 1 all of it, 2 its indices); lds =
// bytes of dynamic LDS
hipError_t launch_forest(const DevModel& d, const JitModule* jm, int ksteps, int waves, int stage, int lds,
                         hipStream_t s);
// packs, plain drop-in: elimination only (rhs/diag from results)
hipError_t launch_forest_elim(const DevModel& d, int lds, hipStream_t s);
hipError_t launch_step_advance(const DevModel& d, int32_t by, hipStream_t s);

// Adjoint (hsh_adjoint.hip), tile layout: one reverse step in three
// launches.  Arrays are [position] (block*64+lane) unless noted.
struct AdjModel {
    DevModel d;           // the model; d.S / d.V are not used
    const double* Vt;     // V at step t
    const double* St;     // slots at step t [nS]
    double* RB;           // rhs of step t
    double* RD;           // diag of step t
    double* LV;           // adjoint of V (in: t+1, out: t)
    double* LS;           // slot adjoints [nS]: gate states (in: t+1, out: t), constants (accumulated)
    double* BB;           // adjoint of rhs
    double* DB;           // adjoint of diag
    double* GT;           // accumulated gradients of ta, tcf, tcb [3][npos]
    double* Dp;           // scratch: eliminated diag, eliminated rhs, solution, adjoint of solution
    double* Bp;
    double* X;
    double* XB;
    long long npos;
    int max_rows;         // rows of the largest tile
};
hipError_t launch_adj_rhs(const AdjModel& a, hipStream_t s);
hipError_t launch_adj_elim(const AdjModel& a, hipStream_t s);
hipError_t launch_adj_local(const AdjModel& a, hipStream_t s);
// LDS bytes per pack position of k_adj_forest: twelve doubles (working
// arrays, gradients, coefficients), parent and observed index, and at most
// one forward entry (two ints), one kid (a short), one backward entry and
// two level bounds per row
constexpr int ADJ_PACK_LDS_PER_POS = 12 * 8 + 2 * 4 + 2 * 4 + 2 + 4 + 2 * 4;
// the pack layout's reverse launch (AdjPack, hsh_devtypes.h); jm may be null
hipError_t launch_adj_forest(const AdjPack& a, const JitModule* jm, int max_pack_pos, hipStream_t s);
hipError_t launch_gather(const double* V, const int* pos, int n, double* out, hipStream_t s);
hipError_t launch_inject(double* LV, const int* pos, int n, const double* g, hipStream_t s);
int batched_lds_bytes(int max_rows);

}  // namespace hsh
