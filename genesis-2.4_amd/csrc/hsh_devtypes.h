// hsh_devtypes.h -- types shared by the host library, the built-in kernels and
// the kernels libhsolve_hip generates and compiles at run time.  Plain C++
// with no includes, so the run-time compiler can take it verbatim.
#pragma once

namespace hsh {

// Lanes per tile: one wave64 lane per neuron.
constexpr int W = 64;

// Device micro-ops.  A compartment's channel program is rewritten into a flat
// list of these; the chip slots it reads are renumbered 0..nslots-1 in the
// order the reference interpreter consumes them.  A program is stored as
// [nslots, template id, micro-ops...].
enum UOp : int {
    U_END = 0,        // results: rhs = Vm + ichan*tbyc, diag = sumg*tbyc + diagterm
    U_NEWVOLT = 1,    // table row + fraction for Vm, with the reference's row cache
    U_CHAN_EK = 2,    // Ek = S[c++]; gmax = Gk = S[c++]
    U_CHAN = 3,       // gmax = Gk = S[c++]  (Ek kept from this compartment)
    U_IPOL1V = 4,     // operands: column, power (+/-1..6); gate state S[c++]
    U_ADDCURR = 5,    // sumg += Gk; ichan += Ek*Gk
    U_SPIKE = 6,      // operand: islot; threshold S[c++]; counter I[islot], reload
                      // I[islot+1], ops index I[islot+2]
};
constexpr int PROG_HEADER = 2;

// Per-block metadata: a block is one row of one 64-neuron tile.
struct BlockMeta {
    int prog;     // offset of the block's program in the program array
    int dslot;    // offset (doubles) of the block's slot area: nslots*64
    int islot;    // offset (ints) of the block's int slots: nislots*64
    int tile;
    int tmpl;     // template (program) id
    int pad0, pad1, pad2;
};

struct TileMeta {
    int block0;   // first block (row 0)
    int rows;     // rows of every neuron in the tile
    int lanes;    // active lanes (neurons)
    int pad;
};

// Device-side argument bundle (passed by value to every kernel).
struct DevModel {
    const int* prog;
    const BlockMeta* blocks;
    const TileMeta* tiles;
    double* S;            // chip slots, [block][slot][lane]
    int* I;               // int slots, [block][islot][lane]
    double* V;            // vm, [block][lane]
    double* R;            // results, [block][2][lane]
    const int* par;       // parent row within the tile, -1 = root; [block][lane]
    const double* ta;     // A[parent][row] (forward numerator), [block][lane]
    const double* tcf;    // A[row][parent] forward, [block][lane]
    const double* tcb;    // A[row][parent] backward, [block][lane]
    const int* pred;      // position of the previous NEWVOLT compartment, -1 none
    const double* tab;    // composite table (xdivs+2)*ncols
    const double* xvals;  // xdivs+2
    int* spk;             // spike ring: [cap][2] (ops index, step)
    int* nspk;            // spike count
    int* step;            // per-tile step counters
    int spk_cap;
    int ncols;
    int xdivs;
    int crank;
    double xmin;
    double invdx;
    int nblocks;
    int ntiles;
    int last_pos;         // device position of the last row (ncompts-1)
    int single;           // ncompts == 1
    int max_slots;        // largest program slot count
    int pad;
};

}  // namespace hsh
