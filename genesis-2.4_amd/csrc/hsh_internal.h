// hsh_internal.h -- shared definitions of libhsolve_hip (host + device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "hsolve_hip.h"

namespace hsh {

// Lanes per tile: one wave64 lane per neuron.
constexpr int W = 64;

// Reference opcodes (genesis/src/hines/hines_defs.h:178-222).
enum RefOp : int32_t {
    COMPT_OP = 100, FCOMPT_OP = 101, LCOMPT_OP = 102, SPIKE_OP = 200, ELMMPUMP_OP = 500,
    CONC_VAL_OP = 1000, NERNST_OP = 2000, GHK_OP = 2010, CHAN_OP = 3000, CHAN_EK_OP = 3001,
    SAME_GHK_OP = 3002, DO_TABGHK_OP = 3003, ADD_CURR_OP = 3200, FLUX_OP = 3300,
    SCALE_FLUX_OP = 3301, STORE_ALLI_OP = 3501, STORE_ALLC_OP = 3502, IPOL1V_OP = 4001,
    IPOL1C_OP = 4002, IPOL2V_OP = 4003, IPOL2C_OP = 4004, SYN2_OP = 4100, SYN3_OP = 4101,
    MGBLOCK_OP = 4200, NEWVOLT_OP = 5100, NEWCONC1_OP = 5110, NEWCONC2_OP = 5111,
    DOMAIN_OP = 5112,
};
// Reference funcs codes (hines_defs.h:119-128).
enum RefFunc : int32_t {
    F_ELIM = 0, F_SET_DIAG = 1, F_SKIP_DIAG = 2, F_FASTSIB = 3, F_COPY = 4, F_SIB = 5,
    F_CALC = 6, F_FINISH = 7,
};

// Device micro-ops.  A compartment's channel program is rewritten into a flat
// list of these; the chip slots it reads are renumbered 0..nslots-1 in the
// order the reference interpreter consumes them, so the device keeps the
// reference's chip cursor discipline with a wave-uniform slot counter.
enum UOp : int32_t {
    U_END = 0,        // results: rhs = Vm + ichan*tbyc, diag = sumg*tbyc + diagterm
    U_NEWVOLT = 1,    // table row + fraction for Vm, with the reference's row cache
    U_CHAN_EK = 2,    // Ek = S[c++]; gmax = Gk = S[c++]
    U_CHAN = 3,       // gmax = Gk = S[c++]  (Ek kept from this compartment)
    U_IPOL1V = 4,     // operands: column, power (+/-1..6); gate state S[c++]
    U_ADDCURR = 5,    // sumg += Gk; ichan += Ek*Gk; Ik = Gk*(Ek-Vm)
    U_SPIKE = 6,      // operand: islot; threshold S[c++]; counter I[islot], reload
                      // I[islot+1], ops index I[islot+2]
};

// Per-block metadata: a block is one row of one 64-neuron tile.
struct BlockMeta {
    int32_t prog;     // offset of the template's code in the program array
    int32_t dslot;    // offset (doubles) of the block's slot area: nslots*64
    int32_t islot;    // offset (ints) of the block's int slots: nislots*64
    int32_t tile;
};

struct TileMeta {
    int32_t block0;   // first block (row 0)
    int32_t rows;     // rows of every neuron in the tile
    int32_t lanes;    // active lanes (neurons)
    int32_t pad;
};

// Device-side argument bundle.
struct DevModel {
    const int32_t* prog;
    const BlockMeta* blocks;
    const TileMeta* tiles;
    double* S;            // chip slots, [block][slot][lane]
    int32_t* I;           // int slots, [block][islot][lane]
    double* V;            // vm, [block][lane]
    double* R;            // results, [block][2][lane]
    const int32_t* par;   // parent row within the tile, -1 = root; [block][lane]
    const double* ta;     // A[parent][row] (forward numerator), [block][lane]
    const double* tcf;    // A[row][parent] forward, [block][lane]
    const double* tcb;    // A[row][parent] backward, [block][lane]
    const int32_t* pred;  // position of the previous NEWVOLT compartment, -1 none
    const double* tab;    // composite table (xdivs+2)*ncols
    const double* xvals;  // xdivs+2
    int32_t* spk;         // spike ring: [cap][2] (ops index, step)
    int32_t* nspk;        // spike count
    int32_t* step;        // per-tile step counters
    int32_t spk_cap;
    int32_t ncols;
    int32_t xdivs;
    int32_t crank;
    double xmin;
    double invdx;
    int32_t nblocks;
    int32_t ntiles;
    int32_t last_pos;     // device position of the last row (ncompts-1)
    int32_t max_slots;    // largest program slot count
    int32_t single;       // ncompts == 1
};

struct Template {
    std::vector<int32_t> code;
    int nslots = 0;
    int nislots = 0;
    int ngates = 0;
};

// Host-side result of analysing a model.
struct Plan {
    std::vector<hsh_refusal> refusals;
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
    std::vector<int32_t> tree_lo;         // per tree (neuron): first row; root = hi
    std::vector<int32_t> tree_hi;
    std::vector<int32_t> pred;            // previous NEWVOLT compartment (global row), -1
    // layout
    std::vector<BlockMeta> blocks;
    std::vector<TileMeta> tiles;
    std::vector<int32_t> prog;
    std::vector<int32_t> comp_pos;        // row -> block*64+lane
    std::vector<int32_t> pos_comp;        // block*64+lane -> row (padding lanes: lane-0 row)
    int64_t nS = 0, nI = 0;
    int max_rows = 0;
    double bytes_per_step = 0;
};

// Analyse ops[] / funcs[] and build the device layout.  Returns false on a
// malformed model (message in err); refusals go into plan.refusals.
bool analyse(const hsh_model& m, Plan& plan, std::string& err);

// kernel launchers (hsh_kernels.hip)
hipError_t launch_channel(const DevModel& d, hipStream_t s);
hipError_t launch_elim(const DevModel& d, int max_rows, hipStream_t s);
// waves == 0 selects the four-wave register-coefficient kernel (rows <= 16)
hipError_t launch_batched(const DevModel& d, int ksteps, int waves, int max_rows, hipStream_t s);
hipError_t launch_iso(const DevModel& d, int ksteps, hipStream_t s);
hipError_t launch_step_advance(const DevModel& d, int32_t by, hipStream_t s);
int batched_lds_bytes(int max_rows);

}  // namespace hsh
