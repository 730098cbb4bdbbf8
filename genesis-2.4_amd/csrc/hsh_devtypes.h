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
                      // I[islot+1], ops index I[islot+2], spike source I[islot+3]
    U_SYN2 = 7,       // operands: stab index k, islot; X = S[c], Y = S[c+1];
                      // counter I[islot], synchan id I[islot+1]; adds the current
    // concentration ops: a pool the compartment owns is a state slot; its
    // flux is a register of the program (pool number p < MAXPOOL)
    U_CONCVAL = 8,    // operand: p; flux[p] = 0
    U_NEWCONC1 = 9,   // operand: slot of the pool's concentration; conc table row
    U_DOMAIN = 10,    // conc = S[c++] * Ik; conc table row
    U_IPOL1C = 11,    // operands: column, power; gate on the conc table row
    U_TABCURR = 12,   // operands: column, table (0 voltage, 1 conc): Ek = A; Gk *= B; adds the current
    U_FLUX = 13,      // operand: p; flux[p] += Ik
    U_SCALEFLUX = 14, // operand: p; flux[p] += S[c++] * Ik
    U_NERNST = 15,    // operands: Cin slot, Cout slot; Ek = log(Cout/Cin) * S[c++]; Gk = S[c++]
    U_GHK = 16,       // operands: Cin slot, Cout slot; GHK slope conductance, Ek; Gk = S[c+2]*slope; c += 3
    U_POOLCA = 17,    // operands: p, conc slot, first of 3 constant slots (base, 1+dt/2tau, dt B):
                      // Ca_concen update
    U_SYN2MG = 18,    // U_SYN2, then the Mg block before the current is added:
                      // X = S[c+2] exp(Vm S[c+3]); Gk *= X / (X + S[c+4])
    U_NEWCONC2 = 19,  // operand: conc slot; table register = conc row, ipol = cipol, then the
                      // second concentration's row
    U_IPOLG = 20,     // operands: which (2: Vm x conc, 3: conc x conc), column, power; 2-D gate
    U_TABGHK = 21,    // operands: column, which (0..3); Ek = A, GHK slope = B (no current)
    U_SAMEGHK = 22,   // Gk = S[c++] * GHK slope
    U_GIV0 = 23,      // operands: Rm slot, islot of the compartment's givals base: leak current
    // U_TABCURR's "which" takes the same four values as U_IPOLG / U_TABGHK
    U_ELMMPUMP = 24,  // operands: conc slot, codiag slot (-1: none); Kd = S[c], I scale = S[c+1],
                      // diagonal scale = S[c+2]: electrogenic pump current and diagonal
    // The compartment's concentration program (its part of concops[]) and
    // its diffusion systems, after its channel ops.  csum is the program's
    // running sum of the concentration it is on; codiag / coresult of a
    // shell are slots (1 and 0 between steps); slot operands are absolute.
    U_CZERO = 25,     // csum = 0
    U_CSTORE = 26,    // operand: coresult slot; S = csum
    U_FLUXCONC = 27,  // operands: first of 2 value slots, flux register (-1: none, 0)
    U_CTAUPUMP = 28,  // operands: first of 2 value slots, codiag slot
    U_TAUPUMP = 29,   // operands: first of 5 value slots, codiag slot (rate from Vm)
    U_MMPUMP = 30,    // operands: first of 2 value slots, conc slot, codiag slot
    U_HILLPUMP = 31,  // operands: first of 2 value slots, conc slot, codiag slot, Hill coefficient
    U_PHILLPUMP = 32, // operands: first of 3 value slots, shell conc slot, its codiag slot, Hill
    U_BUFFER = 33,    // operands: first of 3 value slots, buffer conc slot, pool conc slot
    U_POOL = 34,      // operand: pool conc slot; S += csum
    U_DIFF = 35,      // operands: shells, rows per shell, fixed buffers, first slots of: the
                      // concentrations, covals0 values, buffer rates, codiag per shell,
                      // coresult per shell, elimination scratch (shells*rows*rows), eliminated
                      // right-hand sides (shells*rows)
};
constexpr int MAXPOOL = 4;
constexpr int PROG_HEADER = 2;

// Per-block metadata.  A block is 64 compartments that run the same channel
// program: one row of a 64-neuron tile (lane = neuron), or a run of
// same-program compartments of a pack (lane = compartment).
struct BlockMeta {
    int prog;     // offset of the block's program in the program array
    int dslot;    // offset (doubles) of the block's slot area: nslots*64
    int islot;    // offset (ints) of the block's int slots: nislots*64
    int tile;     // step-counter index: tile t, or ntiles + pack p
    int tmpl;     // template (program) id
    int lanes;    // lanes holding a compartment (the rest pad)
    int pad1, pad2;
};

struct TileMeta {
    int block0;   // first block (row 0)
    int rows;     // rows of every neuron in the tile
    int lanes;    // active lanes (neurons)
    int units;    // 1: each lane is a funcs[] unit (symmetric compartments), solved by
                  // the funcs interpreter; the tile kernels never see these tiles
};

// A pack: whole neurons of any sizes whose compartments fill blocks
// block0..block0+nblk-1, grouped by channel program.  Its Hines elimination
// runs level by level over pack-local positions lp = (block-block0)*64+lane:
//   forward  -- levels of height 1..H: each entry is a parent that pulls its
//               children's contributions in ascending row order;
//   backward -- levels of depth 0..D: roots solve, then children from their
//               parents.
// lvl[lv0 .. lv0+nfl] are the forward level bounds (entry indices relative to
// fe0), lvl[lv0+nfl+1 .. lv0+nfl+1+nbl] the backward bounds (relative to be0).
struct PackMeta {
    int block0, nblk;
    int lv0, nfl, nbl;
    int fe0, nfe;     // forward entries: fent[2e] = parent lp | nkids << 16, fent[2e+1] = first kid
    int ke0, nke;     // kid entries: kidx (lp), kta, ktcf (kid index relative to ke0)
    int be0, nbe;     // backward entries: bkp = lp | (parent lp + 1) << 16, btcb
    int last_lp;      // lp of the reference's last row (ncompts-1), -1 if elsewhere
};

struct EvCtl {
    double tnext;         // SimulationTime() of the next step
    double tnow;          // of the step being run (set by its spike pass)
    double tlast;         // of the last step run
    int err;              // EV_ERR_* bits
    int npost;            // posted events the next post pass applies
    double tstep[2];      // SimulationTime() of step s at [s & 1] (set by its spike pass)
};

// Where the spike pass finds a spike source's values in the layout.
struct SpikeSrc {
    int vpos;             // V position of its compartment
    int spos;             // S position of its threshold
    int cpos;             // I positions of its counter and reload
    int rpos;
    int opidx;            // ops index of the SPIKE_OP
    int out0, out1;       // its out-connections: out_slot / out_sid [out0, out1)
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
    const double* tab;    // composite table (max(xdivs, ydivs)+2)*ncols, zero padded
    long long tab_max;    // largest start offset a 2-D lookup through the table register may use
    const double* xvals;  // xdivs+2
    int* spk;             // spike ring: [cap][2] (ops index, step)
    int* nspk;            // spike count
    int* step;            // step counters: tiles, then packs
    // packs (lane = compartment layout)
    const PackMeta* packs;
    int* pstep;           // step counters of the packs (step + ntiles)
    const int* lvl;
    const int* fent;
    const unsigned short* kidx;
    const double* kta;
    const double* ktcf;
    const int* bkp;
    const double* btcb;
    int npacks;
    // synaptic events
    const double* stab;   // stablist: xconst2, yconst1, yconst2, xconst1(/dt) per type
    // synchans' pending event lists (set when the model has them): the
    // SYN2_OP of each synchan keeps its own list, in the order of
    // newconn/synchan.c:161-358, from the spikes the spike pass of the step
    // decided before the step ran
    int ev;               // 1: lists on; SPIKE_OPs run in the spike pass, not in programs
    // arrivals, double-buffered by step parity so that the spike pass of
    // step s + 1 may run while the synchans of step s still read theirs
    int* narr;            // EVENTs arriving at step s: narr[(s & 1) * ev_nsyn + synchan] (the spike pass
                          // counts, the synchan clears)
    int* arr;             // their in-connection slots, at (s & 1) * ev_nconn + in_off[synchan] + 0..narr-1,
                          // unordered
    int ev_nsyn;          // synchans
    int ev_nconn;         // in-connections (room for the arrivals of one step)
    const int* in_off;    // per synchan: in-connections, by source ops index then connection order
    const int* in_mid;    // per synchan: first in-connection whose source comes after its SYN2_OP
    const int* in_T;      // steps to the hit, (int)(delay/dt + 0.9)
    const double* in_wdt; // (float)weight / dt
    const int* syn_op;    // per synchan: ops index of its SYN2_OP
    const int* lst_off;   // per synchan: first node of its list
    const int* lst_cap;   // nodes it may hold
    int* lst_n;           // pending events
    int* lst_t;           // node time fields
    double* lst_w;        // node weight / dt
    double* tle;          // per synchan: time_last_event
    const double* tnow;   // SimulationTime() of step s: tnow[s & 1] (EvCtl::tstep)
    int* everr;           // error bits (a full list)
    // the spike pass (k_ev_spikes, or the step kernel's look-ahead)
    const SpikeSrc* src;  // spike sources, by position
    const int* src_boff;  // per block: its sources src[src_boff[b] .. src_boff[b+1])
    const int* out_slot;  // out-connections of the sources (SpikeSrc::out0/out1):
    const int* out_sid;   // the in-connection slot and the synchan
    EvCtl* evctl;
    int nsrc;
    int look;             // 1: the step kernel runs the next step's spike pass for its blocks
    double dt;
    // concentration-indexed rows of tab
    double ymin;
    double invdy;
    int ydivs;
    int spk_cap;
    double* giv;          // givals output (nullptr: not produced)
    int chanmode5;        // givals per gmax
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
    int ext;              // some program holds an op beyond the voltage-gated/synaptic set
    // recording of a batched pack launch for the adjoint (null: off): the
    // observed positions' voltages after each step k, trace[k*nobs + j], the
    // state each step starts from, recV[k*rec_npos + pos] and
    // recS[k*rec_nS + slot], and its elimination
    double* trace;
    const int* obs_idx;   // per position: j of an observed row, -1
    int nobs;
    double* recV;
    double* recS;
    double* recB;         // each step's eliminated rhs and diag and its solution,
    double* recD;         // recB[k*rec_npos + pos]
    double* recX;
    long long rec_npos;
    long long rec_nS;
};

// Adjoint on the pack layout: every reverse step of one checkpoint segment in
// one launch, one workgroup per pack, from what a recording forest launch
// kept (DevModel::recV / recS / recB / recD).  Arrays are [position] unless
// noted.
struct AdjPack {
    DevModel d;
    const double* segV;   // [n][npos] state each step of the segment starts from
    const double* segS;   // [n][nS]
    const double* segB;   // [n][npos] each step's eliminated rhs and diag, solution
    const double* segD;
    const double* segX;
    long long npos, nS;
    int t0, n;            // first step of the segment, its steps
    const double* dtrace; // [nsteps][nobs] dL/d(observed V after step t)
    int nobs;
    const int* obs_idx;   // observed row index j, -1
    double* LV;           // adjoint of V, carried from segment to segment
    double* LS;           // slot adjoints [nS]
    double* GT;           // accumulated gradients of ta, tcf, tcb [3][npos]
    const int* ppar;      // pack-local parent position, -1 root, -2 padding
    const double* pta;    // the row's coefficients (0 for roots)
    const double* ptcf;
    const double* ptcb;
};

}  // namespace hsh
