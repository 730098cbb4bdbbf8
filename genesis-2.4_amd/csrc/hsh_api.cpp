// hsh_api.cpp -- C ABI of libhsolve_hip (declared in include/hsolve_hip.h).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <unordered_map>
#include <string>
#include <utility>
#include <vector>

#include "hsh_internal.h"

using namespace hsh;

namespace {

thread_local std::string g_err;

hsh_status fail(hsh_status st, const std::string& msg)
{
    g_err = msg;
    return st;
}

#define HIPCHK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(HSH_ERR_DEVICE, std::string("HIP: ") + #call + " failed: " +  \
                                            hipGetErrorString(e_));                   \
    } while (0)

#define STCHK(call)                         \
    do {                                    \
        hsh_status s_ = (call);             \
        if (s_ != HSH_OK) return s_;        \
    } while (0)

template <class T> hipError_t dalloc(T** p, size_t n)
{
    return hipMalloc((void**)p, std::max<size_t>(1, n) * sizeof(T));
}

template <class T> hipError_t dput(T* dst, const std::vector<T>& src)
{
    if (src.empty()) return hipSuccess;
    return hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}

constexpr int SPIKE_CAP = 1 << 20;
constexpr int GRAPH_MAX_LAUNCHES = 64;
constexpr int MAX_LDS = 160 * 1024;     // gfx950 LDS per workgroup

// Batched kernel classes: contiguous tile ranges (tiles are sorted by rows,
// largest first) that one kernel advances, plus the packs.
enum ClassKind { C_ISO, C_R1, C_R2, C_R4, C_WIDE, C_PLAINLOOP, C_PACKS };

struct KClass {
    int kind;
    int t0, t1;      // tile range (C_PACKS: pack range)
    int rows;        // largest rows in the range
};

}  // namespace

// The merged arrays of a group (hsh_create_group), owned by its handle.
struct GroupArrays {
    std::vector<int32_t> ops, concops;
    std::vector<double> chip, tab, xvals, stab, ravals, vm, conc, concchip, giv;
    hsh::GroupSpec spec;
};

struct hsh_handle {
    hsh_model m{};
    std::unique_ptr<GroupArrays> garr;   // a group's merged arrays (h->m points into them)
    std::vector<hsh_member> members;     // where each solver's values sit in the arrays
    hsh_options opt{};
    Plan plan;
    bool refused = false;
    bool host_only = false;
    int device = 0;
    hipStream_t stream = nullptr;
    // host copies in the reference layout; chip holds the values slots are
    // loaded from: chip[nchips] | conc[nconcs] | concchip[nconcchip]
    std::vector<int32_t> ops;
    std::vector<double> chip;
    // model-constant device buffers
    int32_t* d_prog = nullptr;
    double *d_tab = nullptr, *d_xvals = nullptr, *d_stab = nullptr;
    // synchans' pending event lists (hsh_set_synapses, hsh_post_events),
    // indexed by synchan id (plan.syn_ops) and spike source id
    // (plan.spike_ops)
    bool ev_on = false;
    EvCtl* d_evctl = nullptr;
    int32_t *d_syn_op = nullptr, *d_in_off = nullptr, *d_in_mid = nullptr, *d_in_T = nullptr,
            *d_lst_off = nullptr, *d_lst_cap = nullptr, *d_lst_n = nullptr, *d_lst_t = nullptr, *d_ipos = nullptr,
            *d_narr = nullptr, *d_arr = nullptr, *d_out_slot = nullptr, *d_out_sid = nullptr,
            *d_post_sid = nullptr, *d_post_T = nullptr, *d_post_after = nullptr;
    std::vector<int32_t> out_off;      // per spike source: its out-connections in d_out_slot / d_out_sid
    double *d_in_wdt = nullptr, *d_lst_w = nullptr, *d_tle = nullptr, *d_post_wdt = nullptr;
    SpikeSrc* d_spksrc = nullptr;      // spike sources, sorted by V position
    int32_t* d_src_boff = nullptr;     // per block: its first source in d_spksrc
    int arr_room = 1;                  // arrivals one step's buffer holds (in-connections)
    size_t lst_total = 0;              // list nodes allocated
    int post_cap = 0;                  // posted events the device buffers hold
    struct Post {
        int32_t sid, T, after, seq;
        double wdt;
    };
    std::vector<Post> posts;           // posted since the last step, applied before the next
    std::vector<double> stab_host;     // stablist (message activations are scaled by it)
    InEntry* d_inputs = nullptr;       // staging of hsh_set_inputs / hsh_set_chip entries
    int inputs_cap = 0;
    // host staging of hsh_set_inputs_static, sized at creation
    struct InTagged {
        InEntry e;
        int32_t seq;
    };
    std::vector<InTagged> in_tag;
    std::vector<InEntry> in_host;
    int max_k = 1 << 30;               // batched steps per launch allowed (shortest connection - 1)
    int nconn = 0;
    double t_next = 0.0;               // SimulationTime() of the next step
    long long tab_max = 0;             // largest start offset of a register-based 2-D lookup
    int32_t *d_spk = nullptr, *d_nspk = nullptr;
    double* d_giv = nullptr;           // givals output (models created with givals)
    // layout-dependent device buffers
    BlockMeta* d_blocks = nullptr;
    TileMeta* d_tiles = nullptr;
    PackMeta* d_packs = nullptr;
    double *d_S = nullptr, *d_V = nullptr, *d_R = nullptr, *d_ta = nullptr, *d_tcf = nullptr,
           *d_tcb = nullptr, *d_kta = nullptr, *d_ktcf = nullptr, *d_btcb = nullptr;
    int32_t *d_I = nullptr, *d_par = nullptr, *d_pred = nullptr, *d_step = nullptr, *d_lvl = nullptr,
            *d_fent = nullptr, *d_bkp = nullptr;
    uint16_t* d_kidx = nullptr;
    // funcs[] units (symmetric compartments): units, programs, mutable ravals
    FuncsUnitDev* d_funits = nullptr;
    int32_t* d_ffuncs = nullptr;
    double* d_fwork = nullptr;
    DevModel dm{};
    JitModule jit;
    std::vector<KClass> classes;
    bool plain_loop = false;           // some tile too large for the batched kernels
    int batch_k = 32, batch_waves = 4;
    int pack_waves = 4;
    int pack_stage = 0;       // pack schedule: 0 in memory, 1 all in LDS, 2 its indices in LDS
    bool tuned = false;
    double autotune_s = 0.0;
    // graphs keyed by (K, launches)
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;
    std::map<std::pair<int, int>, hipGraph_t> graph_src;
    // timing
    bool timing = false;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    std::vector<hipEvent_t> evs;      // per plain step: 3 events
    int nev_used = 0;
    bool total_valid = false;
    // chip index -> device S positions (for hsh_set_chip)
    std::vector<int64_t> c2s_off, c2s;
    int host_step = 0;
    // adjoint: checkpoints of the last hsh_adj_forward run, and buffers
    struct Adj {
        int nsteps = 0, ckpt = 0, ncp = 0, nobs = 0;
        bool valid = false;
        size_t npos = 0, nS = 0, nI = 0, nstep = 0;   // sizes the buffers were made for
        int32_t* obs_pos = nullptr;
        double* trace = nullptr;       // [nsteps][nobs]
        double* dtrace = nullptr;
        double *ckV = nullptr, *ckS = nullptr;       // [ncp][...]
        int32_t *ckI = nullptr, *ckStep = nullptr;
        double *segV = nullptr, *segS = nullptr;     // [ckpt][...]
        double *segB = nullptr, *segD = nullptr, *segX = nullptr;   // pack layout: each step's
                                                                    // elimination [ckpt][npos]
        int nbuf = 1;                 // segment buffers (pack layout: two, one being recorded
                                      // while the other is run backwards)
        double *RB = nullptr, *RD = nullptr, *LV = nullptr, *LS = nullptr, *BB = nullptr, *DB = nullptr,
               *GT = nullptr, *Dp = nullptr, *Bp = nullptr, *X = nullptr, *XB = nullptr;
        // pack layout: observed row per position, the tree per position
        bool packs = false;
        int32_t *obs_idx = nullptr, *ppar = nullptr;
        double *pta = nullptr, *ptcf = nullptr, *ptcb = nullptr;
    } adj;
    // the pack adjoint's reverse launches run on a stream of their own, beside
    // the recording launches on `stream`
    hipStream_t adj_stream = nullptr;
    hipEvent_t adj_rec[2] = {nullptr, nullptr}, adj_rev[2] = {nullptr, nullptr};
};

namespace {

void pack_S(hsh_handle* h, std::vector<double>& S)
{
    const Plan& P = h->plan;
    S.assign(P.nS, 0.0);
    for (size_t b = 0; b < P.blocks.size(); b++) {
        const BlockMeta& bm = P.blocks[b];
        for (int l = 0; l < W; l++) {
            int g = P.pos_comp[b * W + l];
            const Template& t = P.templates[P.comp_tmpl[g]];
            const int64_t off = P.comp_dsrc_off[g];
            for (int k = 0; k < t.nslots; k++) {
                int64_t src = P.dsrc[off + k];
                S[(size_t)bm.dslot + (size_t)k * W + l] = src >= 0 ? h->chip[src] : P.synth[-src - 1];
            }
        }
    }
}

void pack_I(hsh_handle* h, std::vector<int32_t>& I)
{
    const Plan& P = h->plan;
    I.assign(std::max<int64_t>(1, P.nI), 0);
    for (size_t b = 0; b < P.blocks.size(); b++) {
        const BlockMeta& bm = P.blocks[b];
        for (int l = 0; l < W; l++) {
            int g = P.pos_comp[b * W + l];
            const Template& t = P.templates[P.comp_tmpl[g]];
            const int64_t off = P.comp_isrc_off[g];
            for (int k = 0; k < t.nislots; k++) {
                int64_t src = P.isrc[off + k];
                I[(size_t)bm.islot + (size_t)k * W + l] = src >= 0 ? h->ops[src] : (int32_t)(-src - 1);
            }
        }
    }
}

void pack_V(hsh_handle* h, const double* vm, std::vector<double>& V)
{
    const Plan& P = h->plan;
    V.assign(P.blocks.size() * W, 0.0);
    for (size_t q = 0; q < V.size(); q++) V[q] = vm[P.pos_comp[q]];
}

// givals around runs that must not show (autotuning, adjoint recomputation)
hsh_status save_givals(hsh_handle* h, std::vector<double>& g)
{
    g.clear();
    if (!h->d_giv) return HSH_OK;
    g.resize((size_t)h->m.ngivals);
    HIPCHK(hipMemcpy(g.data(), h->d_giv, g.size() * sizeof(double), hipMemcpyDeviceToHost));
    return HSH_OK;
}

hsh_status load_givals(hsh_handle* h, const std::vector<double>& g)
{
    if (!h->d_giv || g.empty()) return HSH_OK;
    HIPCHK(hipMemcpy(h->d_giv, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice));
    return HSH_OK;
}

hsh_status upload_state(hsh_handle* h, const double* vm)
{
    std::vector<double> S, V;
    std::vector<int32_t> I;
    pack_S(h, S);
    pack_I(h, I);
    HIPCHK(hipMemcpyAsync(h->d_S, S.data(), S.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_I, I.data(), I.size() * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    if (vm) {
        pack_V(h, vm, V);
        HIPCHK(hipMemcpyAsync(h->d_V, V.data(), V.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return HSH_OK;
}

// Pull the device state back into the handle's host copies of chip[] / ops[].
hsh_status download_chip_ops(hsh_handle* h)
{
    const Plan& P = h->plan;
    std::vector<double> S(P.nS);
    std::vector<int32_t> I(std::max<int64_t>(1, P.nI));
    HIPCHK(hipMemcpyAsync(S.data(), h->d_S, S.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(I.data(), h->d_I, I.size() * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const int n = h->m.ncompts;
    for (int g = 0; g < n; g++) {
        const int pos = P.comp_pos[g];
        const int b = pos / W, l = pos % W;
        const BlockMeta& bm = P.blocks[b];
        const Template& t = P.templates[P.comp_tmpl[g]];
        for (int k = 0; k < t.nslots; k++) {
            int64_t src = P.dsrc[P.comp_dsrc_off[g] + k];
            if (src >= 0) h->chip[src] = S[(size_t)bm.dslot + (size_t)k * W + l];
        }
        for (int k = 0; k < t.nislots; k++) {
            int64_t src = P.isrc[P.comp_isrc_off[g] + k];
            if (src >= 0) h->ops[src] = I[(size_t)bm.islot + (size_t)k * W + l];
        }
    }
    return HSH_OK;
}

// vm[] and results[] in the reference layout.
hsh_status download_vm_results(hsh_handle* h, double* vm, double* results)
{
    const Plan& P = h->plan;
    const size_t npos = P.blocks.size() * W;
    std::vector<double> V(npos), R(2 * npos);
    HIPCHK(hipMemcpy(V.data(), h->d_V, npos * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(R.data(), h->d_R, 2 * npos * sizeof(double), hipMemcpyDeviceToHost));
    for (int g = 0; g < h->m.ncompts; g++) {
        const int pos = P.comp_pos[g];
        const int b = pos / W, l = pos % W;
        if (vm) vm[g] = V[pos];
        if (results) {
            results[2 * g] = R[(size_t)b * 2 * W + l];
            results[2 * g + 1] = R[(size_t)b * 2 * W + W + l];
        }
    }
    return HSH_OK;
}

void free_graphs(hsh_handle* h)
{
    for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.second);
    for (auto& g : h->graph_src) (void)hipGraphDestroy(g.second);
    h->graphs.clear();
    h->graph_src.clear();
}

void free_adj(hsh_handle* h)
{
    auto& A = h->adj;
    void* ptrs[] = {A.obs_pos, A.trace, A.dtrace, A.ckV, A.ckS, A.ckI, A.ckStep, A.segV, A.segS, A.segB,
                    A.segD, A.segX, A.RB, A.RD, A.LV, A.LS, A.BB, A.DB, A.GT, A.Dp, A.Bp, A.X, A.XB, A.obs_idx, A.ppar, A.pta,
                    A.ptcf, A.ptcb};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    A = hsh_handle::Adj();
}

void free_layout(hsh_handle* h)
{
    free_adj(h);
    free_graphs(h);
    void* ptrs[] = {h->d_blocks, h->d_tiles, h->d_packs, h->d_S, h->d_V, h->d_R, h->d_ta, h->d_tcf,
                    h->d_tcb, h->d_kta, h->d_ktcf, h->d_btcb, h->d_I, h->d_par, h->d_pred, h->d_step,
                    h->d_lvl, h->d_fent, h->d_bkp, h->d_kidx, h->d_funits, h->d_ffuncs, h->d_fwork};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    h->d_blocks = nullptr;
    h->d_tiles = nullptr;
    h->d_packs = nullptr;
    h->d_S = h->d_V = h->d_R = h->d_ta = h->d_tcf = h->d_tcb = h->d_kta = h->d_ktcf = h->d_btcb = nullptr;
    h->d_I = h->d_par = h->d_pred = h->d_step = h->d_lvl = h->d_fent = h->d_bkp = nullptr;
    h->d_kidx = nullptr;
    h->d_funits = nullptr;
    h->d_ffuncs = nullptr;
    h->d_fwork = nullptr;
}

void free_events(hsh_handle* h)
{
    void* ptrs[] = {h->d_evctl, h->d_syn_op, h->d_in_off, h->d_in_mid, h->d_in_T, h->d_lst_off, h->d_lst_cap,
                    h->d_lst_n, h->d_lst_t, h->d_ipos, h->d_narr, h->d_arr, h->d_out_slot, h->d_out_sid,
                    h->d_post_sid, h->d_post_T, h->d_post_after, h->d_in_wdt, h->d_lst_w, h->d_tle,
                    h->d_post_wdt, h->d_spksrc, h->d_src_boff};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    h->d_evctl = nullptr;
    h->d_src_boff = nullptr;
    h->d_syn_op = h->d_in_off = h->d_in_mid = h->d_in_T = h->d_lst_off = h->d_lst_cap = nullptr;
    h->d_lst_n = h->d_lst_t = h->d_ipos = h->d_narr = h->d_arr = h->d_out_slot = h->d_out_sid = nullptr;
    h->d_post_sid = h->d_post_T = h->d_post_after = nullptr;
    h->out_off.clear();
    h->d_in_wdt = h->d_lst_w = h->d_tle = h->d_post_wdt = nullptr;
    h->d_spksrc = nullptr;
    h->ev_on = false;
    h->lst_total = 0;
    h->post_cap = 0;
}

void free_device(hsh_handle* h)
{
    free_layout(h);
    void* ptrs[] = {h->d_prog, h->d_tab, h->d_xvals, h->d_stab, h->d_spk, h->d_nspk, h->d_giv, h->d_inputs};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    h->d_inputs = nullptr;
    h->inputs_cap = 0;
    free_events(h);
    h->d_prog = nullptr;
    h->d_tab = h->d_xvals = h->d_stab = nullptr;
    h->d_spk = h->d_nspk = nullptr;
    h->d_giv = nullptr;
    jit_free(h->jit);
    for (hipEvent_t e : h->evs) (void)hipEventDestroy(e);
    h->evs.clear();
    if (h->ev_a) (void)hipEventDestroy(h->ev_a);
    if (h->ev_b) (void)hipEventDestroy(h->ev_b);
    h->ev_a = h->ev_b = nullptr;
    for (int b = 0; b < 2; b++) {
        if (h->adj_rec[b]) (void)hipEventDestroy(h->adj_rec[b]);
        if (h->adj_rev[b]) (void)hipEventDestroy(h->adj_rev[b]);
        h->adj_rec[b] = h->adj_rev[b] = nullptr;
    }
    if (h->adj_stream) (void)hipStreamDestroy(h->adj_stream);
    h->adj_stream = nullptr;
    if (h->stream) (void)hipStreamDestroy(h->stream);
    h->stream = nullptr;
}

// Tile classes for the batched step (tiles are sorted by rows, largest first).
void compute_classes(hsh_handle* h)
{
    const Plan& P = h->plan;
    h->classes.clear();
    h->plain_loop = false;
    auto kind_of = [](int rows) {
        if (rows == 1) return (int)C_ISO;
        if (rows <= 4) return (int)C_R1;
        if (rows <= 8) return (int)C_R2;
        if (rows <= 16) return (int)C_R4;
        if (rows <= 64) return (int)C_WIDE;
        return (int)C_PLAINLOOP;
    };
    // funcs[] units run their own interpreter, so a handle with them steps
    // plain (two launches and the units' one)
    if (!P.funits.empty()) h->plain_loop = true;
    for (int t = 0; t < P.ntiles_regular; t++) {
        const int k = kind_of(P.tiles[t].rows);
        if (k == C_PLAINLOOP) h->plain_loop = true;
        if (!h->classes.empty() && h->classes.back().kind == k) {
            h->classes.back().t1 = t + 1;
            h->classes.back().rows = std::max(h->classes.back().rows, P.tiles[t].rows);
        } else {
            h->classes.push_back(KClass{k, t, t + 1, P.tiles[t].rows});
        }
    }
    if (!P.packs.empty()) h->classes.push_back(KClass{C_PACKS, 0, (int)P.packs.size(), 0});
}

// Does any program hold an op outside the interpreter's basic set (END,
// NEWVOLT, CHAN, CHAN_EK, IPOL1V, ADDCURR, SPIKE, SYN2, GIV0)?  Operand
// counts follow the micro-op table of hsh_devtypes.h; an op not named here
// counts as extended, so the full interpreter runs it.
bool programs_extended(const Plan& P)
{
    for (const Template& t : P.templates) {
        const std::vector<int32_t>& code = t.code;
        for (size_t pc = 0; pc < code.size();) {
            const int op = code[pc++];
            if (op == U_END) break;
            if (op == U_IPOL1V || op == U_SYN2 || op == U_GIV0) {
                pc += 2;
            } else if (op == U_SPIKE) {
                pc += 1;
            } else if (op != U_NEWVOLT && op != U_CHAN && op != U_CHAN_EK && op != U_ADDCURR) {
                return true;
            }
        }
    }
    return false;
}


void set_devmodel(hsh_handle* h)
{
    const Plan& P = h->plan;
    DevModel& d = h->dm;
    d = DevModel{};
    d.prog = h->d_prog;
    d.blocks = h->d_blocks;
    d.tiles = h->d_tiles;
    d.S = h->d_S;
    d.I = h->d_I;
    d.V = h->d_V;
    d.R = h->d_R;
    d.par = h->d_par;
    d.ta = h->d_ta;
    d.tcf = h->d_tcf;
    d.tcb = h->d_tcb;
    d.pred = h->d_pred;
    d.tab = h->d_tab;
    d.xvals = h->d_xvals;
    d.spk = h->d_spk;
    d.nspk = h->d_nspk;
    d.step = h->d_step;
    d.packs = h->d_packs;
    d.pstep = h->d_step + P.tiles.size();
    d.lvl = h->d_lvl;
    d.fent = h->d_fent;
    d.kidx = h->d_kidx;
    d.kta = h->d_kta;
    d.ktcf = h->d_ktcf;
    d.bkp = h->d_bkp;
    d.btcb = h->d_btcb;
    d.npacks = (int32_t)P.packs.size();
    d.stab = h->d_stab;
    d.ev = h->ev_on ? 1 : 0;
    d.narr = h->d_narr;
    d.arr = h->d_arr;
    d.ev_nsyn = (int32_t)P.syn_ops.size();
    d.ev_nconn = h->arr_room;
    d.in_off = h->d_in_off;
    d.in_mid = h->d_in_mid;
    d.in_T = h->d_in_T;
    d.in_wdt = h->d_in_wdt;
    d.syn_op = h->d_syn_op;
    d.lst_off = h->d_lst_off;
    d.lst_cap = h->d_lst_cap;
    d.lst_n = h->d_lst_n;
    d.lst_t = h->d_lst_t;
    d.lst_w = h->d_lst_w;
    d.tle = h->d_tle;
    d.tnow = h->d_evctl ? h->d_evctl->tstep : nullptr;
    d.everr = h->d_evctl ? &h->d_evctl->err : nullptr;
    d.src = h->d_spksrc;
    d.src_boff = h->d_src_boff;
    d.out_slot = h->d_out_slot;
    d.out_sid = h->d_out_sid;
    d.evctl = h->d_evctl;
    d.nsrc = (int32_t)P.spike_ops.size();
    d.look = 0;
    d.dt = h->m.dt;
    d.spk_cap = SPIKE_CAP;
    d.giv = P.giv ? h->d_giv : nullptr;
    d.chanmode5 = P.chanmode5 ? 1 : 0;
    d.ncols = h->m.ncols;
    d.xdivs = h->m.xdivs;
    d.crank = (h->m.method == HSH_CRANK);
    d.xmin = h->m.xmin;
    d.invdx = h->m.invdx;
    d.tab_max = h->tab_max;
    d.ymin = h->m.ymin;
    d.invdy = h->m.invdy;
    d.ydivs = h->m.ydivs;
    d.nblocks = (int32_t)P.blocks.size();
    d.ntiles = (int32_t)P.tiles.size();
    d.last_pos = P.comp_pos[h->m.ncompts - 1];
    d.single = (h->m.ncompts == 1);
    d.max_slots = 0;
    for (const Template& t : P.templates) d.max_slots = std::max(d.max_slots, t.nslots);
    d.ext = programs_extended(P) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Synchans' pending event lists (hsh_events.hip, syn2_op in hsh_device.h)

// Position in the layout's int slots of each synchan's SYN2_OP counter, and
// where the spike pass finds each spike source's values.
void ev_positions(const hsh_handle* h, std::vector<int32_t>& ipos, std::vector<SpikeSrc>& src)
{
    const Plan& P = h->plan;
    std::unordered_map<int64_t, int32_t> sid_of_cnt, cnt_of_spk, rel_of_spk;
    for (size_t i = 0; i < P.syn_ops.size(); i++) sid_of_cnt[(int64_t)P.syn_ops[i] + 2] = (int32_t)i;
    for (size_t i = 0; i < P.spike_ops.size(); i++) {
        cnt_of_spk[(int64_t)P.spike_ops[i] + 1] = (int32_t)i;
        rel_of_spk[(int64_t)P.spike_ops[i] + 2] = (int32_t)i;
    }
    ipos.assign(P.syn_ops.size(), -1);
    src.assign(P.spike_ops.size(), SpikeSrc{-1, -1, -1, -1, -1, 0, 0, 0});
    for (int g = 0; g < h->m.ncompts; g++) {
        const Template& t = P.templates[P.comp_tmpl[g]];
        const int pos = P.comp_pos[g];
        const BlockMeta& bm = P.blocks[pos / W];
        for (int k = 0; k < t.nislots; k++) {
            const int64_t srcix = P.isrc[P.comp_isrc_off[g] + k];
            if (srcix < 0) continue;
            const int ip = bm.islot + k * W + pos % W;
            auto a = sid_of_cnt.find(srcix);
            if (a != sid_of_cnt.end()) ipos[a->second] = ip;
            auto b = cnt_of_spk.find(srcix);
            if (b != cnt_of_spk.end()) src[b->second].cpos = ip;
            auto c = rel_of_spk.find(srcix);
            if (c != rel_of_spk.end()) src[c->second].rpos = ip;
        }
    }
    for (size_t i = 0; i < P.spike_ops.size(); i++) {
        const int64_t ci = P.spike_thresh[i];
        src[i].vpos = P.comp_pos[P.spike_comp[i]];
        if (ci >= 0 && ci + 1 < (int64_t)h->c2s_off.size() && h->c2s_off[ci] < h->c2s_off[ci + 1])
            src[i].spos = (int)h->c2s[h->c2s_off[ci]];
        src[i].opidx = P.spike_ops[i];
        src[i].out0 = h->out_off.empty() ? 0 : h->out_off[i];
        src[i].out1 = h->out_off.empty() ? 0 : h->out_off[i + 1];
    }
}

// After a (re)layout: the positions the spike and post passes use.
hsh_status ev_layout(hsh_handle* h)
{
    if (!h->ev_on) return HSH_OK;
    std::vector<int32_t> ipos;
    std::vector<SpikeSrc> src;
    ev_positions(h, ipos, src);
    for (const SpikeSrc& q : src)
        if (q.vpos < 0 || q.spos < 0 || q.cpos < 0 || q.rpos < 0)
            return fail(HSH_ERR_ARG, "hsh: a SPIKE_OP's values are not in the device layout");
    for (int32_t q : ipos)
        if (q < 0) return fail(HSH_ERR_ARG, "hsh: a SYN2_OP counter is not in the device layout");
    // by position, so that each block's sources are one range (the step
    // kernel's look-ahead spike pass takes its blocks' ranges)
    std::stable_sort(src.begin(), src.end(), [](const SpikeSrc& a, const SpikeSrc& b) { return a.vpos < b.vpos; });
    const size_t nb = h->plan.blocks.size();
    std::vector<int32_t> boff(nb + 1, 0);
    for (const SpikeSrc& q : src) boff[q.vpos / W + 1]++;
    for (size_t b = 0; b < nb; b++) boff[b + 1] += boff[b];
    if (h->d_src_boff) HIPCHK(hipFree(h->d_src_boff));
    h->d_src_boff = nullptr;
    HIPCHK(dalloc(&h->d_src_boff, nb + 1));
    HIPCHK(dput(h->d_src_boff, boff));
    HIPCHK(dput(h->d_ipos, ipos));
    HIPCHK(dput(h->d_spksrc, src));
    set_devmodel(h);
    return HSH_OK;
}

// Sticky errors of the lists.
hsh_status ev_check(hsh_handle* h)
{
    if (!h->ev_on) return HSH_OK;
    int32_t err = 0;
    HIPCHK(hipMemcpy(&err, &h->d_evctl->err, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (err & EV_ERR_FULL)
        return fail(HSH_ERR_LIMIT, "hsh: a synchan's pending event list ran out of room; events were lost");
    return HSH_OK;
}

// The lists' host image: what ev_rebuild keeps across a new connection table
// and what the autotuner saves and restores (counters live in ops[]).
struct EvImage {
    std::vector<int32_t> n, off, t;
    std::vector<double> w, tle;
    EvCtl ctl{};
};

hsh_status ev_download(hsh_handle* h, EvImage& im)
{
    const size_t ns = h->plan.syn_ops.size();
    HIPCHK(hipStreamSynchronize(h->stream));
    im.n.resize(ns);
    im.off.resize(ns);
    im.tle.resize(ns);
    im.t.resize(h->lst_total);
    im.w.resize(h->lst_total);
    HIPCHK(hipMemcpy(im.n.data(), h->d_lst_n, ns * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(im.off.data(), h->d_lst_off, ns * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(im.tle.data(), h->d_tle, ns * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(im.t.data(), h->d_lst_t, h->lst_total * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(im.w.data(), h->d_lst_w, h->lst_total * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&im.ctl, h->d_evctl, sizeof(EvCtl), hipMemcpyDeviceToHost));
    return HSH_OK;
}

// Put back an image taken with the same list offsets.
hsh_status ev_upload(hsh_handle* h, const EvImage& im)
{
    const size_t ns = h->plan.syn_ops.size();
    HIPCHK(hipMemcpy(h->d_lst_n, im.n.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_tle, im.tle.data(), ns * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_lst_t, im.t.data(), h->lst_total * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_lst_w, im.w.data(), h->lst_total * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_evctl, &im.ctl, sizeof(EvCtl), hipMemcpyHostToDevice));
    return HSH_OK;
}

// (Re)build the list buffers for a connection table: in-connections of each
// synchan sorted by the source's ops index (the order a step's EVENTs reach
// it) then connection order; list capacities from the most events a
// connection can hold pending (one per reload interval of its source over
// its delay, plus two) with room for 32 more (posted events, early fires
// shifting nothing later).  Pending events and time_last_event carry over; a
// first call starts every list empty.  Counters stay in ops[].
hsh_status ev_rebuild(hsh_handle* h, const std::vector<int32_t>& csrc, const std::vector<int32_t>& cdst,
                      const std::vector<int32_t>& cT, const std::vector<double>& cwdt)
{
    const Plan& P = h->plan;
    const size_t ns = P.syn_ops.size(), nsrc = P.spike_ops.size();
    const size_t n = csrc.size();
    EvImage old;
    const bool had = h->ev_on;
    if (had) STCHK(ev_download(h, old));
    STCHK(download_chip_ops(h));
    std::vector<int32_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = (int32_t)i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        if (cdst[a] != cdst[b]) return cdst[a] < cdst[b];
        return P.spike_ops[csrc[a]] < P.spike_ops[csrc[b]];
    });
    std::vector<int32_t> in_off(ns + 1, 0), in_srcop(std::max<size_t>(1, n)), in_T(std::max<size_t>(1, n)),
        slot_of(n), in_mid(ns);
    std::vector<double> in_wdt(std::max<size_t>(1, n));
    std::vector<int64_t> cap(ns, 32);
    for (size_t q = 0; q < n; q++) {
        const int32_t i = order[q];
        in_off[cdst[i] + 1]++;
        slot_of[i] = (int32_t)q;
        in_srcop[q] = P.spike_ops[csrc[i]];
        in_T[q] = cT[i];
        in_wdt[q] = cwdt[i];
        const int32_t reload = std::max(1, h->ops[P.spike_ops[csrc[i]] + 2]);
        cap[cdst[i]] += (cT[i] + 2 + reload - 1) / reload;
    }
    for (size_t i = 0; i < ns; i++) in_off[i + 1] += in_off[i];
    for (size_t i = 0; i < ns; i++) {
        int32_t q = in_off[i];
        while (q < in_off[i + 1] && in_srcop[q] < P.syn_ops[i]) q++;
        in_mid[i] = q;
    }
    // out-connections by source: the in-connection slot and synchan a spike flags
    std::vector<int32_t> out_off(nsrc + 1, 0), out_slot(std::max<size_t>(1, n)), out_sid(std::max<size_t>(1, n));
    for (size_t i = 0; i < n; i++) out_off[csrc[i] + 1]++;
    for (size_t i = 0; i < nsrc; i++) out_off[i + 1] += out_off[i];
    {
        std::vector<int32_t> cur(out_off.begin(), out_off.end() - 1);
        for (size_t i = 0; i < n; i++) {
            const int32_t o = cur[csrc[i]]++;
            out_slot[o] = slot_of[i];
            out_sid[o] = cdst[i];
        }
    }
    std::vector<int32_t> lst_off(ns), lst_cap(ns), lst_n(ns, 0), syn_op(ns);
    std::vector<double> tle(ns, 0.0);
    int64_t total = 0;
    for (size_t i = 0; i < ns; i++) {
        if (had) cap[i] = std::max<int64_t>(cap[i], old.n[i] + 32);
        lst_off[i] = (int32_t)total;
        lst_cap[i] = (int32_t)cap[i];
        total += cap[i];
        if (total > ((int64_t)1 << 30))
            return fail(HSH_ERR_LIMIT, "hsh_set_synapses: pending-list capacity too large");
        syn_op[i] = P.syn_ops[i];
    }
    std::vector<int32_t> lt((size_t)std::max<int64_t>(1, total), 0);
    std::vector<double> lw((size_t)std::max<int64_t>(1, total), 0.0);
    EvCtl ctl{};
    if (had) {
        for (size_t i = 0; i < ns; i++) {
            lst_n[i] = old.n[i];
            tle[i] = old.tle[i];
            for (int k = 0; k < old.n[i]; k++) {
                lt[lst_off[i] + k] = old.t[old.off[i] + k];
                lw[lst_off[i] + k] = old.w[old.off[i] + k];
            }
        }
        ctl = old.ctl;
    } else {
        ctl.tnext = h->t_next;
        ctl.tnow = h->t_next - h->m.dt;
        ctl.tlast = ctl.tnow - h->m.dt;
    }
    free_events(h);
    const size_t ntot = lt.size();
    HIPCHK(dalloc(&h->d_evctl, 1));
    HIPCHK(dalloc(&h->d_syn_op, ns));
    HIPCHK(dalloc(&h->d_in_off, ns + 1));
    HIPCHK(dalloc(&h->d_in_mid, ns));
    HIPCHK(dalloc(&h->d_out_slot, out_slot.size()));
    HIPCHK(dalloc(&h->d_out_sid, out_sid.size()));
    // arrivals: two buffers, by step parity
    h->arr_room = (int)std::max<size_t>(1, n);
    HIPCHK(dalloc(&h->d_arr, 2 * (size_t)h->arr_room));
    HIPCHK(dalloc(&h->d_narr, 2 * std::max<size_t>(1, ns)));
    HIPCHK(dalloc(&h->d_in_T, in_T.size()));
    HIPCHK(dalloc(&h->d_in_wdt, in_wdt.size()));
    HIPCHK(dalloc(&h->d_lst_off, ns));
    HIPCHK(dalloc(&h->d_lst_cap, ns));
    HIPCHK(dalloc(&h->d_lst_n, ns));
    HIPCHK(dalloc(&h->d_lst_t, ntot));
    HIPCHK(dalloc(&h->d_lst_w, ntot));
    HIPCHK(dalloc(&h->d_tle, ns));
    HIPCHK(dalloc(&h->d_ipos, ns));
    HIPCHK(dalloc(&h->d_spksrc, std::max<size_t>(1, nsrc)));
    HIPCHK(hipMemcpy(h->d_evctl, &ctl, sizeof(EvCtl), hipMemcpyHostToDevice));
    HIPCHK(dput(h->d_syn_op, syn_op));
    HIPCHK(dput(h->d_in_off, in_off));
    HIPCHK(dput(h->d_in_mid, in_mid));
    HIPCHK(dput(h->d_out_slot, out_slot));
    HIPCHK(dput(h->d_out_sid, out_sid));
    HIPCHK(hipMemset(h->d_narr, 0, 2 * std::max<size_t>(1, ns) * sizeof(int32_t)));
    HIPCHK(dput(h->d_in_T, in_T));
    HIPCHK(dput(h->d_in_wdt, in_wdt));
    HIPCHK(dput(h->d_lst_off, lst_off));
    HIPCHK(dput(h->d_lst_cap, lst_cap));
    HIPCHK(dput(h->d_lst_n, lst_n));
    HIPCHK(dput(h->d_lst_t, lt));
    HIPCHK(dput(h->d_lst_w, lw));
    HIPCHK(dput(h->d_tle, tle));
    h->out_off = out_off;
    h->lst_total = ntot;
    h->ev_on = true;
    STCHK(ev_layout(h));
    set_devmodel(h);
    return HSH_OK;
}

// Room for n posted events in the device buffers.
hsh_status ev_post_room(hsh_handle* h, int n)
{
    if (n <= h->post_cap) return HSH_OK;
    int cap = std::max(1024, h->post_cap);
    while (cap < n) cap *= 2;
    void* ptrs[] = {h->d_post_sid, h->d_post_T, h->d_post_after, h->d_post_wdt};
    for (void* p : ptrs)
        if (p) HIPCHK(hipFree(p));
    HIPCHK(dalloc(&h->d_post_sid, (size_t)cap));
    HIPCHK(dalloc(&h->d_post_T, (size_t)cap));
    HIPCHK(dalloc(&h->d_post_after, (size_t)cap));
    HIPCHK(dalloc(&h->d_post_wdt, (size_t)cap));
    h->post_cap = cap;
    return HSH_OK;
}

// Apply the events posted since the last step.
hsh_status ev_flush(hsh_handle* h)
{
    if (!h->ev_on || h->posts.empty()) return HSH_OK;
    std::vector<hsh_handle::Post>& ps = h->posts;
    std::stable_sort(ps.begin(), ps.end(), [](const hsh_handle::Post& a, const hsh_handle::Post& b) {
        return a.sid != b.sid ? a.sid < b.sid : a.seq < b.seq;
    });
    const int n = (int)ps.size();
    STCHK(ev_post_room(h, n));
    std::vector<int32_t> sid(n), T(n), after(n);
    std::vector<double> w(n);
    for (int i = 0; i < n; i++) {
        sid[i] = ps[i].sid;
        T[i] = ps[i].T;
        after[i] = ps[i].after;
        w[i] = ps[i].wdt;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(dput(h->d_post_sid, sid));
    HIPCHK(dput(h->d_post_T, T));
    HIPCHK(dput(h->d_post_after, after));
    HIPCHK(dput(h->d_post_wdt, w));
    HIPCHK(hipMemcpy(&h->d_evctl->npost, &n, sizeof(int32_t), hipMemcpyHostToDevice));
    const EvPost p{h->d_post_sid, h->d_post_T, h->d_post_wdt, h->d_post_after, h->d_ipos,
                   (int)h->plan.syn_ops.size()};
    HIPCHK(launch_ev_post(h->dm, p, h->d_evctl, h->stream));
    const int zero = 0;
    HIPCHK(hipMemcpyAsync(&h->d_evctl->npost, &zero, sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    ps.clear();
    return HSH_OK;
}

// The step's spike pass (lists on): every SPIKE_OP before the step's kernels.
hsh_status ev_spikes(hsh_handle* h)
{
    if (!h->ev_on) return HSH_OK;
    HIPCHK(launch_ev_spikes(h->dm, h->stream));
    return HSH_OK;
}

// Allocate and fill the layout-dependent buffers from the plan and the given
// state (vm, results in the reference layout; chip/ops from the host copies).
hsh_status alloc_layout(hsh_handle* h, const double* vm, const double* results)
{
    const Plan& P = h->plan;
    const size_t npos = P.blocks.size() * W;
    const size_t ncnt = P.tiles.size() + P.packs.size();
    HIPCHK(dalloc(&h->d_blocks, P.blocks.size()));
    HIPCHK(dalloc(&h->d_tiles, P.tiles.size()));
    HIPCHK(dalloc(&h->d_packs, P.packs.size()));
    HIPCHK(dalloc(&h->d_S, (size_t)P.nS));
    HIPCHK(dalloc(&h->d_I, (size_t)std::max<int64_t>(1, P.nI)));
    HIPCHK(dalloc(&h->d_V, npos));
    HIPCHK(dalloc(&h->d_R, 2 * npos));
    HIPCHK(dalloc(&h->d_par, npos));
    HIPCHK(dalloc(&h->d_ta, npos));
    HIPCHK(dalloc(&h->d_tcf, npos));
    HIPCHK(dalloc(&h->d_tcb, npos));
    HIPCHK(dalloc(&h->d_pred, npos));
    HIPCHK(dalloc(&h->d_step, ncnt));
    HIPCHK(dalloc(&h->d_lvl, P.lvl.size()));
    HIPCHK(dalloc(&h->d_fent, P.fent.size()));
    HIPCHK(dalloc(&h->d_kidx, P.kidx.size()));
    HIPCHK(dalloc(&h->d_kta, P.kta.size()));
    HIPCHK(dalloc(&h->d_ktcf, P.ktcf.size()));
    HIPCHK(dalloc(&h->d_bkp, P.bkp.size()));
    HIPCHK(dalloc(&h->d_btcb, P.btcb.size()));
    if (!P.funits.empty()) {
        std::vector<FuncsUnitDev> fu(P.funits.size());
        for (size_t u = 0; u < P.funits.size(); u++) {
            fu[u].tile = P.funit_tile[u];
            fu[u].lane = P.funit_lane[u];
            fu[u].nn = P.funits[u].nn;
            fu[u].pad = 0;
            fu[u].f0 = P.funits[u].f0;
            fu[u].q0 = P.funits[u].q0;
        }
        HIPCHK(dalloc(&h->d_funits, fu.size()));
        HIPCHK(dalloc(&h->d_ffuncs, P.ffuncs.size()));
        HIPCHK(dalloc(&h->d_fwork, P.fravals.size()));
        HIPCHK(dput(h->d_funits, fu));
        HIPCHK(dput(h->d_ffuncs, P.ffuncs));
        HIPCHK(dput(h->d_fwork, P.fravals));
    }
    // tree and predecessor arrays of the tile layout (row within the tile)
    std::vector<int32_t> par(npos), pred(npos);
    std::vector<double> ta(npos), tcf(npos), tcb(npos);
    {
        std::vector<int32_t> lo_of(h->m.ncompts);
        for (size_t t = 0; t < P.tree_lo.size(); t++)
            for (int r = P.tree_lo[t]; r <= P.tree_hi[t]; r++) lo_of[r] = P.tree_lo[t];
        for (size_t q = 0; q < npos; q++) {
            int g = P.pos_comp[q];
            int pg = P.parent[g];
            par[q] = pg >= 0 ? pg - lo_of[g] : -1;
            ta[q] = P.ta[g];
            tcf[q] = P.tcf[g];
            tcb[q] = P.tcb[g];
            const bool real = (P.comp_pos[g] == (int)q);
            pred[q] = (real && P.pred[g] >= 0) ? P.comp_pos[P.pred[g]] : -1;
        }
    }
    HIPCHK(dput(h->d_blocks, P.blocks));
    HIPCHK(dput(h->d_tiles, P.tiles));
    HIPCHK(dput(h->d_packs, P.packs));
    HIPCHK(dput(h->d_par, par));
    HIPCHK(dput(h->d_pred, pred));
    HIPCHK(dput(h->d_ta, ta));
    HIPCHK(dput(h->d_tcf, tcf));
    HIPCHK(dput(h->d_tcb, tcb));
    HIPCHK(dput(h->d_lvl, P.lvl));
    HIPCHK(dput(h->d_fent, P.fent));
    HIPCHK(dput(h->d_kidx, P.kidx));
    HIPCHK(dput(h->d_kta, P.kta));
    HIPCHK(dput(h->d_ktcf, P.ktcf));
    HIPCHK(dput(h->d_bkp, P.bkp));
    HIPCHK(dput(h->d_btcb, P.btcb));
    std::vector<int32_t> steps(ncnt, h->host_step);
    HIPCHK(dput(h->d_step, steps));
    std::vector<double> R(2 * npos, 0.0);
    if (results)
        for (size_t q = 0; q < npos; q++) {
            const int g = P.pos_comp[q];
            const size_t b = q / W, l = q % W;
            R[b * 2 * W + l] = results[2 * g];
            R[b * 2 * W + W + l] = results[2 * g + 1];
        }
    HIPCHK(dput(h->d_R, R));
    STCHK(upload_state(h, vm));
    // reverse map chip index -> S positions
    {
        const int nchips = (int)h->chip.size();
        std::vector<int64_t> cnt(nchips + 1, 0);
        for (int g = 0; g < h->m.ncompts; g++) {
            const Template& t = P.templates[P.comp_tmpl[g]];
            for (int k = 0; k < t.nslots; k++) {
                int64_t src = P.dsrc[P.comp_dsrc_off[g] + k];
                if (src >= 0) cnt[src + 1]++;
            }
        }
        for (int i = 0; i < nchips; i++) cnt[i + 1] += cnt[i];
        h->c2s_off = cnt;
        h->c2s.assign(cnt[nchips], 0);
        std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
        for (int g = 0; g < h->m.ncompts; g++) {
            const Template& t = P.templates[P.comp_tmpl[g]];
            const int pos = P.comp_pos[g];
            const BlockMeta& bm = P.blocks[pos / W];
            for (int k = 0; k < t.nslots; k++) {
                int64_t src = P.dsrc[P.comp_dsrc_off[g] + k];
                if (src >= 0) h->c2s[cur[src]++] = (int64_t)bm.dslot + (int64_t)k * W + pos % W;
            }
        }
    }
    set_devmodel(h);
    compute_classes(h);
    return ev_layout(h);
}

// Move the device state to a new layout.
hsh_status relayout(hsh_handle* h, const LayoutOpts& o)
{
    HIPCHK(hipStreamSynchronize(h->stream));
    const int n = h->m.ncompts;
    std::vector<double> vm(n), res(2 * (size_t)n);
    STCHK(download_vm_results(h, vm.data(), res.data()));
    STCHK(download_chip_ops(h));
    std::string err;
    if (!build_layout(h->plan, o, err)) return fail(HSH_ERR_ARG, "layout: " + err);
    free_layout(h);
    return alloc_layout(h, vm.data(), res.data());
}

int64_t pack_lds_of(const Plan& P, int stage)
{
    return stage == 1 ? P.pack_lds_stage : stage == 2 ? P.pack_lds_stage_idx : P.pack_lds;
}

int pack_lds_bytes(const hsh_handle* h, int stage) { return (int)pack_lds_of(h->plan, stage); }

bool stage_fits(const hsh_handle* h, int stage) { return pack_lds_of(h->plan, stage) <= MAX_LDS; }

DevModel tile_view(const DevModel& d, const KClass& c)
{
    DevModel v = d;
    v.tiles = d.tiles + c.t0;
    v.step = d.step + c.t0;
    v.ntiles = c.t1 - c.t0;
    return v;
}

hsh_status launch_class(hsh_handle* h, const KClass& c, int K, hipStream_t s, bool look = false)
{
    if (c.kind == C_PACKS) {
        HIPCHK(launch_forest(h->dm, &h->jit, K, h->pack_waves, h->pack_stage, pack_lds_bytes(h, h->pack_stage), s));
        return HSH_OK;
    }
    DevModel v = tile_view(h->dm, c);
    v.look = look ? 1 : 0;
    if (c.kind == C_ISO) {
        HIPCHK(launch_iso(v, &h->jit, K, s));
    } else {
        const int waves = (c.kind == C_WIDE) ? std::max(1, h->batch_waves) : 0;
        HIPCHK(launch_batched(v, &h->jit, K, waves, c.rows, s));
    }
    return HSH_OK;
}

// With the lists on and every tile one row (one kernel class, no packs), the
// step kernel runs the next step's spike pass for its own blocks as soon as
// it has solved them (body_iso): a run of n steps is one spike launch and n
// step launches, the last of which leaves the next step's pass to the next
// run (which may follow a change of the state, the inputs or the posted
// events).  Other layouts launch the spike pass before every step.
bool lookahead_ok(const hsh_handle* h)
{
    return h->ev_on && h->classes.size() == 1 && h->classes[0].kind == C_ISO && h->plan.packs.empty() &&
           h->plan.funits.empty() && h->d_src_boff != nullptr;
}

// One batched launch of K steps: the kernel of every class, in turn, on the
// handle's stream (K = 1 when the lists are on; `first` and `last` place the
// launch in a run of launches).
hsh_status batched_launch(hsh_handle* h, int K, bool first, bool last)
{
    const bool look = lookahead_ok(h);
    if (!look || first) STCHK(ev_spikes(h));
    for (const KClass& c : h->classes) STCHK(launch_class(h, c, K, h->stream, look && !last));
    return HSH_OK;
}

hsh_status run_batched(hsh_handle* h, int nsteps, int* launches)
{
    const int K = std::max(1, std::min(h->batch_k, h->max_k));
    int full = nsteps / K;
    const int rem = nsteps % K;
    while (full > 0) {
        const int M = std::min(full, GRAPH_MAX_LAUNCHES);
        auto key = std::make_pair(K, M);
        auto it = h->graphs.find(key);
        if (it == h->graphs.end()) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            hsh_status st = HSH_OK;
            for (int i = 0; i < M && st == HSH_OK; i++) st = batched_launch(h, K, i == 0, i == M - 1);
            hipError_t ce = hipStreamEndCapture(h->stream, &g);
            if (st != HSH_OK) return st;
            if (ce != hipSuccess) return fail(HSH_ERR_DEVICE, std::string("HIP: graph capture failed: ") + hipGetErrorString(ce));
            hipGraphExec_t ex = nullptr;
            HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
            h->graph_src[key] = g;
            it = h->graphs.emplace(key, ex).first;
        }
        HIPCHK(hipGraphLaunch(it->second, h->stream));
        (*launches)++;
        full -= M;
    }
    if (rem > 0) {
        STCHK(batched_launch(h, rem, true, true));
        (*launches)++;
    }
    return HSH_OK;
}

// One plain step: channel programs of every block, then the eliminations.
hsh_status plain_step(hsh_handle* h, bool channels_only)
{
    if (h->timing) {
        while ((int)h->evs.size() < h->nev_used + 3) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            h->evs.push_back(e);
        }
        HIPCHK(hipEventRecord(h->evs[h->nev_used], h->stream));
    }
    STCHK(ev_spikes(h));
    HIPCHK(launch_channel(h->dm, &h->jit, h->stream));
    if (h->timing) HIPCHK(hipEventRecord(h->evs[h->nev_used + 1], h->stream));
    if (channels_only) {
        HIPCHK(launch_step_advance(h->dm, 1, h->stream));
    } else {
        if (h->plan.ntiles_regular > 0) {
            DevModel dr = h->dm;
            dr.ntiles = h->plan.ntiles_regular;
            HIPCHK(launch_elim(dr, h->plan.max_rows, h->stream));
        }
        if (h->dm.npacks > 0) HIPCHK(launch_forest_elim(h->dm, pack_lds_bytes(h, false), h->stream));
        if (!h->plan.funits.empty())
            HIPCHK(launch_funcs(h->dm, h->d_funits, (int)h->plan.funits.size(), h->d_ffuncs, h->d_fwork,
                                h->stream));
    }
    if (h->timing) {
        HIPCHK(hipEventRecord(h->evs[h->nev_used + 2], h->stream));
        h->nev_used += 3;
    }
    return HSH_OK;
}

// Device time per step of `steps` batched steps (graph replays included).
hsh_status time_batched(hsh_handle* h, int steps, double* ms_per_step)
{
    int launches = 0;
    STCHK(run_batched(h, steps, &launches));        // warm (and capture)
    HIPCHK(hipEventRecord(h->ev_a, h->stream));
    STCHK(run_batched(h, steps, &launches));
    HIPCHK(hipEventRecord(h->ev_b, h->stream));
    HIPCHK(hipEventSynchronize(h->ev_b));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, h->ev_a, h->ev_b));
    *ms_per_step = (double)t / steps;
    return HSH_OK;
}

// The autotuner: layout (tile_max_rows, pack_cap), pack workgroup shape
// (waves, schedule in LDS) and K, each swept on this model with the others
// held, fastest kept.  Options the caller fixed are not swept.  Only layouts
// that one kernel runs are candidates (all trees in tiles of one class, or
// all in packs); a layout that mixes classes runs only as configured, unless
// the library is built with HSH_AUTOTUNE_MIXED=1 (the host-sanitizer build
// of tests/hostsan sweeps them).  The state is saved first and restored
// after (spikes raised while tuning are dropped).
#ifndef HSH_AUTOTUNE_MIXED
#define HSH_AUTOTUNE_MIXED 0
#endif
hsh_status autotune(hsh_handle* h)
{
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipStreamSynchronize(h->stream));
    const int n = h->m.ncompts;
    std::vector<double> vm0(n), res0(2 * (size_t)n);
    STCHK(download_vm_results(h, vm0.data(), res0.data()));
    STCHK(download_chip_ops(h));
    const std::vector<double> chip0 = h->chip;
    const std::vector<int32_t> ops0 = h->ops;
    int32_t nspk0 = 0;
    HIPCHK(hipMemcpy(&nspk0, h->d_nspk, sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<double> giv0;
    STCHK(save_givals(h, giv0));
    // pending synaptic events: trials fire and add them
    STCHK(ev_flush(h));
    EvImage ev0;
    if (h->ev_on) STCHK(ev_download(h, ev0));
    const hsh_options& o = h->opt;
    const Plan& P = h->plan;
    // HSH_AUTOTUNE_LOG=1 prints every timed candidate on stderr
    const char* logenv = std::getenv("HSH_AUTOTUNE_LOG");
    const bool log = logenv && logenv[0] && logenv[0] != '0';
    const int K0 = std::min(o.batch_k > 0 ? o.batch_k : 32, h->max_k);
    h->batch_k = K0;
    // steps per trial: about 2e6 compartment-steps, at least 2 launches
    const int trial = std::max(2 * K0, std::min(4096, (int)(2e6 / std::max(1, n)) / K0 * K0));

    std::vector<int> tmr_c, cap_c, waves_c, stage_c;
    if (o.tile_max_rows != 0) tmr_c = {o.tile_max_rows < 0 ? 0 : o.tile_max_rows};
    else {
        tmr_c = {16};
        if (P.max_tree_rows > 16) tmr_c.push_back(64);
        tmr_c.push_back(0);
    }
    if (o.pack_cap > 0) cap_c = {o.pack_cap};
    else {
        cap_c = {512, 1024, 2048};
        // pack sizes whose packs all fit the GPU at once: k packs per CU,
        // bound by the 28 bytes of LDS per position (with room for block
        // padding), so no workgroup waits for a second round
        int ncu = 0, ldscu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess) ncu = 0;
        if (hipDeviceGetAttribute(&ldscu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, h->device) !=
                hipSuccess || ldscu <= 0)
            ldscu = MAX_LDS;
        for (int k = 4; k <= 7 && ncu > 0; k++) {
            const int64_t cap = (int64_t)((double)n * 1.08 / ((double)ncu * k)) + 1;
            if (cap >= std::max(64, P.max_tree_rows) && cap * 28 * k <= ldscu && cap <= 8192 &&
                std::find(cap_c.begin(), cap_c.end(), (int)cap) == cap_c.end())
                cap_c.push_back((int)cap);
        }
    }
    waves_c = o.pack_waves > 0 ? std::vector<int>{o.pack_waves} : std::vector<int>{4, 8};
    stage_c = o.pack_stage > 0 ? std::vector<int>{o.pack_stage - 1} : std::vector<int>{0, 1, 2};

    struct Best {
        double ms = 1e300;
        LayoutOpts lo;
        int waves = 4;
        int stage = 0;
    } best;
    for (int tmr : tmr_c) {
        for (size_t ci = 0; ci < cap_c.size(); ci++) {
            LayoutOpts lo;
            lo.tile_max_rows = tmr;
            lo.pack_cap = cap_c[ci];
            STCHK(relayout(h, lo));
            // pack size matters only when some tree went to a pack
            const bool any_pack = !h->plan.packs.empty();
            if (h->plain_loop) break;        // layouts with too-large tiles are not batched
            if (!HSH_AUTOTUNE_MIXED && h->classes.size() != 1) break;
            for (int wv : (any_pack ? waves_c : std::vector<int>{waves_c[0]})) {
                for (int st : (any_pack ? stage_c : std::vector<int>{0})) {
                    if (st && !stage_fits(h, st)) continue;
                    if (h->plan.pack_lds > MAX_LDS) continue;
                    h->pack_waves = wv;
                    h->pack_stage = st;
                    free_graphs(h);
                    double ms = 0;
                    STCHK(time_batched(h, trial, &ms));
                    if (log)
                        std::fprintf(stderr, "hsh_autotune: tile_max_rows %d pack_cap %d waves %d stage %d: %.2f us/step (%d tiles, %d packs)\n",
                                     tmr, cap_c[ci], wv, st, ms * 1e3, h->dm.ntiles, h->dm.npacks);
                    if (ms < best.ms) {
                        best.ms = ms;
                        best.lo = lo;
                        best.waves = wv;
                        best.stage = st;
                    }
                }
            }
            if (!any_pack) break;
        }
    }
    if (best.ms < 1e300) {
        STCHK(relayout(h, best.lo));
        h->pack_waves = best.waves;
        h->pack_stage = best.stage;
        if (o.batch_k <= 0) {
            double bk = 1e300;
            int kbest = K0;
            std::vector<int> ks = {1, 8, 32, 128};
            if (h->max_k < 128) ks.push_back(h->max_k);     // the longest launch connections allow
            for (int K : ks) {
                if (K > h->max_k) continue;
                h->batch_k = K;
                free_graphs(h);
                double ms = 0;
                STCHK(time_batched(h, std::max(2 * K, trial / K * K), &ms));
                if (log) std::fprintf(stderr, "hsh_autotune: K %d: %.2f us/step\n", K, ms * 1e3);
                if (ms < bk) {
                    bk = ms;
                    kbest = K;
                }
            }
            h->batch_k = kbest;
            free_graphs(h);
        }
    }
    // restore the state
    h->chip = chip0;
    h->ops = ops0;
    free_layout(h);
    STCHK(alloc_layout(h, vm0.data(), res0.data()));
    HIPCHK(hipMemcpy(h->d_nspk, &nspk0, sizeof(int32_t), hipMemcpyHostToDevice));
    STCHK(load_givals(h, giv0));
    if (h->ev_on) STCHK(ev_upload(h, ev0));
    h->tuned = true;
    h->autotune_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return HSH_OK;
}


// ---------------------------------------------------------------------------
// Adjoint support

constexpr int ADJ_MAXG = 3;     // gates per channel of the reverse program

bool adj_supported(const Plan& P, std::string& why)
{
    if (!P.funits.empty()) {
        why = "symmetric compartments (sibling elimination ops)";
        return false;
    }
    for (size_t t = 0; t < P.templates.size(); t++) {
        const std::vector<int32_t>& code = P.templates[t].code;
        int ng = 0;
        for (size_t pc = 0; pc < code.size();) {
            const int op = code[pc++];
            if (op == U_IPOL1V) {
                pc += 2;
                if (++ng > ADJ_MAXG) {
                    why = "a channel with more than 3 gates";
                    return false;
                }
            } else if (op == U_CHAN || op == U_CHAN_EK) {
                ng = 0;
            } else if (op == U_SPIKE) {
                pc += 1;
            } else if (op == U_SYN2 || op == U_SYN2MG) {
                why = "synaptic channels";
                return false;
            } else if (op >= U_CONCVAL) {
                why = "concentration-dependent ops, 2-D tables or GHK";
                return false;
            } else if (op == U_END) {
                break;
            }
        }
    }
    return true;
}

// The adjoint runs on one layout class: all tiles (three launches per
// reverse step) or all packs (one launch per checkpoint segment).  A layout
// that mixes them moves every tree to packs; packs whose reverse-sweep LDS
// does not fit a workgroup are made smaller.
hsh_status ensure_adj_layout(hsh_handle* h)
{
    if (h->plan.packs.empty()) return HSH_OK;
    LayoutOpts lo = h->plan.layout;
    if (!h->plan.tiles.empty()) {
        lo.tile_max_rows = 0;
        STCHK(relayout(h, lo));
    }
    while ((int64_t)ADJ_PACK_LDS_PER_POS * h->plan.max_pack_pos + 16 > MAX_LDS) {
        if (lo.pack_cap <= std::max(64, h->plan.max_tree_rows))
            return fail(HSH_ERR_LIMIT, "hsh_adj_forward: a tree too large for the adjoint's LDS arrays");
        lo.pack_cap = std::max(std::max(64, h->plan.max_tree_rows), lo.pack_cap / 2);
        STCHK(relayout(h, lo));
    }
    return HSH_OK;
}

// Per-position arrays of the pack adjoint: observed row, pack-local parent
// (-1 root, -2 padding) and the row's coefficients.
hsh_status adj_pack_arrays(hsh_handle* h, int nobs, const int32_t* obs)
{
    auto& A = h->adj;
    const Plan& P = h->plan;
    const size_t npos = A.npos;
    std::vector<int32_t> oi(npos, -1), par(npos, -2);
    std::vector<double> ta(npos, 0.0), tf(npos, 0.0), tb(npos, 0.0);
    for (int j = 0; j < nobs; j++) oi[P.comp_pos[obs[j]]] = j;
    for (int g = 0; g < h->m.ncompts; g++) {
        const int pos = P.comp_pos[g];
        const int pg = P.parent[g];
        if (pg < 0) {
            par[pos] = -1;
            continue;
        }
        par[pos] = P.comp_pos[pg];
        ta[pos] = P.ta[g];
        tf[pos] = P.tcf[g];
        tb[pos] = P.tcb[g];
    }
    // parents as pack-local positions (whole neurons sit in one pack)
    for (const PackMeta& pm : P.packs) {
        const int lo = pm.block0 * W, hi = (pm.block0 + pm.nblk) * W;
        for (int q = lo; q < hi; q++)
            if (par[q] >= 0) {
                if (par[q] < lo || par[q] >= hi) return fail(HSH_ERR_STATE, "hsh_adj_forward: a tree split over packs");
                par[q] -= lo;
            }
    }
    HIPCHK(dalloc(&A.obs_idx, npos));
    HIPCHK(dalloc(&A.ppar, npos));
    HIPCHK(dalloc(&A.pta, npos));
    HIPCHK(dalloc(&A.ptcf, npos));
    HIPCHK(dalloc(&A.ptcb, npos));
    HIPCHK(dput(A.obs_idx, oi));
    HIPCHK(dput(A.ppar, par));
    HIPCHK(dput(A.pta, ta));
    HIPCHK(dput(A.ptcf, tf));
    HIPCHK(dput(A.ptcb, tb));
    return HSH_OK;
}

hsh_status adj_alloc(hsh_handle* h, int nsteps, int ckpt, int nobs)
{
    auto& A = h->adj;
    const Plan& P = h->plan;
    const size_t npos = P.blocks.size() * W, nS = (size_t)std::max<int64_t>(1, P.nS);
    const size_t nI = (size_t)std::max<int64_t>(1, P.nI), nstep = P.tiles.size() + P.packs.size();
    const int ncp = (nsteps + ckpt - 1) / ckpt + 1;
    free_adj(h);
    A.nsteps = nsteps;
    A.ckpt = ckpt;
    A.ncp = ncp;
    A.nobs = nobs;
    A.npos = npos;
    A.nS = nS;
    A.nI = nI;
    A.nstep = std::max<size_t>(1, nstep);
    HIPCHK(dalloc(&A.obs_pos, (size_t)nobs));
    HIPCHK(dalloc(&A.trace, (size_t)nsteps * nobs));
    HIPCHK(dalloc(&A.dtrace, (size_t)nsteps * nobs));
    HIPCHK(dalloc(&A.ckV, (size_t)ncp * npos));
    HIPCHK(dalloc(&A.ckS, (size_t)ncp * nS));
    HIPCHK(dalloc(&A.ckI, (size_t)ncp * nI));
    HIPCHK(dalloc(&A.ckStep, (size_t)ncp * A.nstep));
    A.nbuf = P.packs.empty() ? 1 : 2;
    HIPCHK(dalloc(&A.segV, (size_t)A.nbuf * ckpt * npos));
    HIPCHK(dalloc(&A.segS, (size_t)A.nbuf * ckpt * nS));
    if (!P.packs.empty()) {
        // padding positions are never written: zeros there
        for (double** b : {&A.segB, &A.segD, &A.segX}) {
            HIPCHK(dalloc(b, (size_t)A.nbuf * ckpt * npos));
            HIPCHK(hipMemset(*b, 0, (size_t)A.nbuf * ckpt * npos * sizeof(double)));
        }
    }
    double** bufs[] = {&A.RB, &A.RD, &A.LV, &A.BB, &A.DB, &A.Dp, &A.Bp, &A.X, &A.XB};
    for (double** b : bufs) HIPCHK(dalloc(b, npos));
    HIPCHK(dalloc(&A.LS, nS));
    HIPCHK(dalloc(&A.GT, 3 * npos));
    return HSH_OK;
}

hsh_status adj_save(hsh_handle* h, int i)
{
    auto& A = h->adj;
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    HIPCHK(hipMemcpyAsync(A.ckV + (size_t)i * A.npos, h->d_V, A.npos * sizeof(double), k, h->stream));
    HIPCHK(hipMemcpyAsync(A.ckS + (size_t)i * A.nS, h->d_S, (size_t)h->plan.nS * sizeof(double), k, h->stream));
    HIPCHK(hipMemcpyAsync(A.ckI + (size_t)i * A.nI, h->d_I, (size_t)std::max<int64_t>(1, h->plan.nI) * sizeof(int32_t),
                          k, h->stream));
    HIPCHK(hipMemcpyAsync(A.ckStep + (size_t)i * A.nstep, h->d_step,
                          (h->plan.tiles.size() + h->plan.packs.size()) * sizeof(int32_t), k, h->stream));
    return HSH_OK;
}

hsh_status adj_load(hsh_handle* h, int i)
{
    auto& A = h->adj;
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    HIPCHK(hipMemcpyAsync(h->d_V, A.ckV + (size_t)i * A.npos, A.npos * sizeof(double), k, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_S, A.ckS + (size_t)i * A.nS, (size_t)h->plan.nS * sizeof(double), k, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_I, A.ckI + (size_t)i * A.nI, (size_t)std::max<int64_t>(1, h->plan.nI) * sizeof(int32_t),
                          k, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_step, A.ckStep + (size_t)i * A.nstep,
                          (h->plan.tiles.size() + h->plan.packs.size()) * sizeof(int32_t), k, h->stream));
    return HSH_OK;
}

}  // namespace

extern "C" {

const char* hsh_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace {

hsh_status check_model(const hsh_model* m)
{
    if (m->ncompts <= 0 || m->nops <= 0 || m->nchips <= 0 || !m->ops || !m->chip || !m->vm ||
        !m->funcs || m->nfuncs <= 0)
        return fail(HSH_ERR_ARG, "hsh_create: empty model or missing ops/chip/vm/funcs");
    if (m->ncols > 0 && (!m->tablist || !m->xvals || m->xdivs <= 0))
        return fail(HSH_ERR_ARG, "hsh_create: ncols > 0 but no table");
    if (m->chanmode != 4 && m->chanmode != 5)
        return fail(HSH_ERR_ARG, "hsh_create: chanmode must be 4 or 5");
    if (m->method != HSH_BEULER && m->method != HSH_CRANK)
        return fail(HSH_ERR_ARG, "hsh_create: method must be HSH_BEULER or HSH_CRANK");
    if (m->nconcs < 0 || (m->nconcs > 0 && !m->conc) || m->nconcops < 0 ||
        (m->nconcops > 0 && !m->concops) || m->nconcchip < 0 || (m->nconcchip > 0 && !m->concchip) ||
        m->ydivs < 0)
        return fail(HSH_ERR_ARG, "hsh_create: inconsistent concentration arrays");
    if (m->ngivals < 0 || (m->ngivals > 0 && !m->givals))
        return fail(HSH_ERR_ARG, "hsh_create: ngivals > 0 but no givals");
    if (m->ndiffs < 0 || m->ncovals0 < 0 || m->nbufvals < 0 ||
        (m->ndiffs > 0 && (!m->nshells || !m->nshrows || !m->nshfix || !m->dconc || !m->fbufval ||
                           !m->fdiffs || (m->ncovals0 > 0 && !m->covals0) || (m->nbufvals > 0 && !m->bufvals))))
        return fail(HSH_ERR_ARG, "hsh_create: inconsistent diffusion arrays");
    return HSH_OK;
}

hsh_status create_impl(const hsh_model* m, const hsh_options* opt, hsh_handle** out, std::unique_ptr<GroupArrays> ga,
                       std::vector<hsh_member> members)
{
    STCHK(check_model(m));
    if (opt && (opt->pack_waves < 0 || opt->pack_waves > 8 || opt->pack_stage < 0 || opt->pack_stage > 3 ||
                opt->pack_cap < 0 || opt->batch_k < 0 || opt->batch_waves < 0))
        return fail(HSH_ERR_ARG, "hsh_create: option out of range");
    hsh_handle* h = new hsh_handle();
    h->m = *m;
    h->garr = std::move(ga);
    h->members = std::move(members);
    if (opt) h->opt = *opt;
    h->ops.assign(m->ops, m->ops + m->nops);
    h->chip.assign(m->chip, m->chip + m->nchips);
    if (m->nstab > 0 && m->stablist) h->stab_host.assign(m->stablist, m->stablist + 4 * (size_t)m->nstab);
    if (m->nconcs > 0) h->chip.insert(h->chip.end(), m->conc, m->conc + m->nconcs);
    if (m->nconcchip > 0) h->chip.insert(h->chip.end(), m->concchip, m->concchip + m->nconcchip);
    if (m->ngivals > 0 && m->givals) h->chip.insert(h->chip.end(), m->givals, m->givals + m->ngivals);
    if (m->ndiffs > 0) {
        if (m->ncovals0 > 0) h->chip.insert(h->chip.end(), m->covals0, m->covals0 + m->ncovals0);
        if (m->nbufvals > 0) h->chip.insert(h->chip.end(), m->bufvals, m->bufvals + m->nbufvals);
    }
    std::string err;
    if (!analyse(*m, h->plan, err, h->garr ? &h->garr->spec : nullptr)) {
        delete h;
        return fail(HSH_ERR_ARG, "hsh_create: " + err);
    }
    if (!h->plan.refusals.empty()) {
        h->refused = true;
        *out = h;
        char buf[160];
        const hsh_refusal& r = h->plan.refusals[0];
        std::snprintf(buf, sizeof buf,
                      "hsh_create: %zu refusal(s); first: compartment %d, code %d at index %d, reason %d",
                      h->plan.refusals.size(), r.comp, r.opcode, r.op_index, r.reason);
        g_err = buf;
        return HSH_REFUSED;
    }
    // initial layout from the options (defaults where not given)
    const hsh_options& o = h->opt;
    LayoutOpts lo;
    if (o.tile_max_rows != 0) lo.tile_max_rows = o.tile_max_rows < 0 ? 0 : o.tile_max_rows;
    if (o.pack_cap > 0) lo.pack_cap = o.pack_cap;
    if (!build_layout(h->plan, lo, err)) {
        delete h;
        return fail(HSH_ERR_ARG, "hsh_create: " + err);
    }
    if (o.batch_k > 0) h->batch_k = o.batch_k;
    if (o.pack_waves > 0) h->pack_waves = o.pack_waves;
    if (o.pack_stage > 0) h->pack_stage = o.pack_stage - 1;
    if (!stage_fits(h, h->pack_stage)) h->pack_stage = 0;
    if (h->plan.pack_lds > MAX_LDS) {
        delete h;
        return fail(HSH_ERR_ARG, "hsh_create: a pack needs more LDS than a workgroup has");
    }
    // waves per tile of the LDS kernel for trees of 17..64 rows
    if (o.batch_waves > 0) h->batch_waves = o.batch_waves;
    else h->batch_waves = std::min(16, std::max(1, h->plan.max_rows / 4));
    if (o.flags & HSH_F_HOST_ONLY) {
        h->host_only = true;
        compute_classes(h);
        *out = h;
        return HSH_OK;
    }
    h->device = o.device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        delete h;
        return fail(HSH_ERR_DEVICE, "hsh_create: no HIP device");
    }
    if (h->device < 0 || h->device >= ndev) {
        delete h;
        return fail(HSH_ERR_ARG, "hsh_create: device ordinal out of range");
    }
    auto cleanup_fail = [&](hsh_status st) {
        free_device(h);
        delete h;
        return st;
    };
#define HIPCHK_C(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return cleanup_fail(fail(HSH_ERR_DEVICE, std::string("HIP: ") + #call +         \
                                                         " failed: " + hipGetErrorString(e_))); \
    } while (0)
    HIPCHK_C(hipSetDevice(h->device));
    // hsh_set_inputs_static: the static staging array's address on this
    // device, looked up once per device with the device current
    HIPCHK_C(inputs_static_setup(h->device));
    h->in_tag.resize(HSH_INPUTS_STATIC_MAX);
    h->in_host.resize(HSH_INPUTS_STATIC_MAX);
    HIPCHK_C(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK_C(hipEventCreate(&h->ev_a));
    HIPCHK_C(hipEventCreate(&h->ev_b));
    const Plan& P = h->plan;
    HIPCHK_C(dalloc(&h->d_prog, P.prog.size()));
    HIPCHK_C(dalloc(&h->d_spk, 2 * (size_t)SPIKE_CAP));
    HIPCHK_C(dalloc(&h->d_nspk, 1));
    // voltage rows and concentration rows share the table
    const size_t trows = (size_t)std::max(m->xdivs, m->ydivs) + 2;
    const size_t ntab = (m->ncols > 0) ? trows * m->ncols : 1;
    // zero padding past the table: a 2-D lookup at the top concentration row
    // reads up to 2*ydivs+3 values past its row (with weight 0)
    const size_t tpad = 2 * (size_t)std::max(0, m->ydivs) + 4 + (size_t)std::max(0, m->ncols);
    const size_t nx = (m->ncols > 0) ? (size_t)(m->xdivs + 2) : 1;
    HIPCHK_C(dalloc(&h->d_tab, ntab + tpad));
    HIPCHK_C(hipMemset(h->d_tab, 0, (ntab + tpad) * sizeof(double)));
    h->tab_max = (long long)(ntab + tpad) - std::max(0, m->ncols) - 4;
    HIPCHK_C(dalloc(&h->d_xvals, nx));
    HIPCHK_C(dalloc(&h->d_stab, (size_t)std::max(1, 4 * m->nstab)));
    if (m->nstab > 0 && m->stablist)
        HIPCHK_C(hipMemcpy(h->d_stab, m->stablist, 4 * (size_t)m->nstab * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK_C(dput(h->d_prog, P.prog));
    if (m->ncols > 0) {
        HIPCHK_C(hipMemcpy(h->d_tab, m->tablist, ntab * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK_C(hipMemcpy(h->d_xvals, m->xvals, nx * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK_C(hipMemset(h->d_nspk, 0, sizeof(int32_t)));
    if (P.giv) {
        HIPCHK_C(dalloc(&h->d_giv, (size_t)m->ngivals));
        HIPCHK_C(hipMemcpy(h->d_giv, m->givals, (size_t)m->ngivals * sizeof(double), hipMemcpyHostToDevice));
    }
#undef HIPCHK_C
    {
        hsh_status st = alloc_layout(h, m->vm, nullptr);
        if (st != HSH_OK) return cleanup_fail(st);
    }
    // kernels specialised to this model's channel programs; on failure the
    // built-in interpreter kernels run (message kept in hsh_last_error)
    const char* nojit = std::getenv("HSH_NO_JIT");
    if (!(o.flags & HSH_F_NO_JIT) && !(nojit && nojit[0] == '1')) {
        std::string jerr;
        if (!jit_build(P, m->ncols, h->jit, jerr)) {
            std::fprintf(stderr, "libhsolve_hip: specialised kernels unavailable (%s); using the interpreter kernels\n",
                         jerr.c_str());
            g_err = jerr;
        }
    }
    *out = h;
    return HSH_OK;
}

bool same_bits(const double* a, const double* b, size_t n)
{
    return n == 0 || std::memcmp(a, b, n * sizeof(double)) == 0;
}

}  // namespace

extern "C" {

hsh_status hsh_create(const hsh_model* m, const hsh_options* opt, hsh_handle** out)
{
    if (!m || !out) return fail(HSH_ERR_ARG, "hsh_create: null argument");
    *out = nullptr;
    std::vector<hsh_member> one(1);
    one[0] = hsh_member{0, m->ncompts, 0, m->nchips, 0, std::max(0, m->nops - 1), 0, std::max(0, m->nconcs), 0,
                        std::max(0, m->ngivals), 0, 0};
    return create_impl(m, opt, out, nullptr, std::move(one));
}

hsh_status hsh_create_group(const hsh_model* const* ms, int32_t n, const hsh_options* opt, hsh_handle** out)
{
    if (!ms || !out || n <= 0) return fail(HSH_ERR_ARG, "hsh_create_group: null argument or no members");
    *out = nullptr;
    for (int32_t i = 0; i < n; i++) {
        if (!ms[i]) return fail(HSH_ERR_ARG, "hsh_create_group: null member");
        STCHK(check_model(ms[i]));
        // members merge Ca_concen pools only: the concentration program of
        // shells, buffers, pumps and pools is owned per compartment of one
        // solver and is not offset across members
        bool conc_prog = ms[i]->ndiffs > 0;
        for (int32_t q = 0; q < ms[i]->nconcops && !conc_prog; q += 2) {
            const int32_t c = ms[i]->concops[q];
            if (c == LCACONC_OP) break;
            if (c != CACONC_OP) conc_prog = true;
        }
        if (conc_prog)
            return fail(HSH_ERR_ARG, "hsh_create_group: a member with diffusion shells, buffers, pumps or "
                                     "concpools (create it as a solver of its own)");
    }
    const hsh_model& m0 = *ms[0];
    const bool giv = m0.ngivals > 0;
    const hsh_model* my = nullptr;               // a member with concentration tables
    const hsh_model* mx = nullptr;               // a member with a gate table
    for (int32_t i = 0; i < n; i++) {
        if (ms[i]->ydivs > 0 && !my) my = ms[i];
        if (ms[i]->ncols > 0 && !mx) mx = ms[i];
    }
    for (int32_t i = 1; i < n; i++) {
        const hsh_model& m = *ms[i];
        // one step, one integration method for all members
        if (m.dt != m0.dt || m.method != m0.method || m.chanmode != m0.chanmode || m.calcmode != m0.calcmode ||
            (m.ngivals > 0) != giv)
            return fail(HSH_ERR_ARG, "hsh_create_group: members differ in dt, method, chanmode, calcmode or givals");
    }
    for (int32_t i = 0; i < n; i++) {
        const hsh_model& m = *ms[i];
        // one voltage axis for every member with tables
        if (m.ncols > 0 && (m.xdivs != mx->xdivs || m.xmin != mx->xmin || m.invdx != mx->invdx ||
                            !same_bits(m.xvals, mx->xvals, (size_t)m.xdivs + 2)))
            return fail(HSH_ERR_ARG, "hsh_create_group: members differ in the voltage table axis");
    }
    for (int32_t i = 0; i < n; i++)
        if (ms[i]->ydivs > 0 && (ms[i]->ydivs != my->ydivs || ms[i]->ymin != my->ymin || ms[i]->invdy != my->invdy))
            return fail(HSH_ERR_ARG, "hsh_create_group: members differ in the concentration table axis");
    auto ga = std::make_unique<GroupArrays>();
    GroupArrays& A = *ga;
    hsh::GroupSpec& G = A.spec;
    std::vector<hsh_member> mem(n);
    // tables: members with the same columns share them
    const int32_t ydivs = my ? my->ydivs : 0;
    const int32_t xdivs = mx ? mx->xdivs : m0.xdivs;
    const size_t trows = (size_t)std::max(xdivs, ydivs) + 2;
    auto rows_of = [](const hsh_model& m) { return (size_t)std::max(m.xdivs, m.ydivs) + 2; };
    std::vector<std::pair<const hsh_model*, int32_t>> tab_seen, stab_seen;
    std::vector<int32_t> col0(n, 0), stab0(n, 0);
    int32_t ncols = 0, nstab = 0;
    for (int32_t i = 0; i < n; i++) {
        const hsh_model& m = *ms[i];
        if (m.ncols > 0) {
            col0[i] = -1;
            for (auto& t : tab_seen)
                if (t.first->ncols == m.ncols && rows_of(*t.first) == rows_of(m) &&
                    same_bits(t.first->tablist, m.tablist, rows_of(m) * m.ncols)) {
                    col0[i] = t.second;
                    break;
                }
            if (col0[i] < 0) {
                col0[i] = ncols;
                tab_seen.emplace_back(&m, ncols);
                ncols += m.ncols;
            }
        }
        if (m.nstab > 0 && m.stablist) {
            stab0[i] = -1;
            for (auto& t : stab_seen)
                if (t.first->nstab == m.nstab && same_bits(t.first->stablist, m.stablist, 4 * (size_t)m.nstab)) {
                    stab0[i] = t.second;
                    break;
                }
            if (stab0[i] < 0) {
                stab0[i] = 4 * nstab;
                stab_seen.emplace_back(&m, 4 * nstab);
                nstab += m.nstab;
            }
        }
    }
    if (ncols > 0) {
        A.tab.assign(trows * ncols, 0.0);
        for (auto& t : tab_seen) {
            // a member with fewer rows repeats its last (guard) row
            const size_t mr = rows_of(*t.first);
            for (size_t r = 0; r < trows; r++)
                std::memcpy(&A.tab[r * ncols + t.second], t.first->tablist + std::min(r, mr - 1) * t.first->ncols,
                            (size_t)t.first->ncols * sizeof(double));
        }
        A.xvals.assign(mx->xvals, mx->xvals + xdivs + 2);
    }
    for (auto& t : stab_seen) A.stab.insert(A.stab.end(), t.first->stablist, t.first->stablist + 4 * t.first->nstab);
    int32_t rows = 0, nconcs = 0;
    for (int32_t i = 0; i < n; i++) {
        const hsh_model& m = *ms[i];
        if (m.ops[m.nops - 1] != LCOMPT_OP) return fail(HSH_ERR_ARG, "hsh_create_group: a member's ops[] lacks LCOMPT_OP");
        hsh_member& e = mem[i];
        e.row0 = rows; e.nrows = m.ncompts;
        e.chip0 = (int32_t)A.chip.size(); e.nchips = m.nchips;
        e.op0 = (int32_t)A.ops.size(); e.nops = m.nops - 1;
        e.conc0 = nconcs; e.nconcs = std::max(0, m.nconcs);
        e.giv0 = (int32_t)A.giv.size(); e.ngivals = giv ? m.ngivals : 0;
        e.stab0 = stab0[i]; e.pad = 0;
        A.ops.insert(A.ops.end(), m.ops, m.ops + m.nops - 1);
        A.chip.insert(A.chip.end(), m.chip, m.chip + m.nchips);
        A.vm.insert(A.vm.end(), m.vm, m.vm + m.ncompts);
        if (m.nconcs > 0) A.conc.insert(A.conc.end(), m.conc, m.conc + m.nconcs);
        if (m.nconcchip > 0) A.concchip.insert(A.concchip.end(), m.concchip, m.concchip + m.nconcchip);
        if (giv) A.giv.insert(A.giv.end(), m.givals, m.givals + m.ngivals);
        // Ca_concen pools of every member in one program, pool indices offset
        for (int32_t q = 0; q + 1 < m.nconcops; q += 2) {
            if (m.concops[q] == LCACONC_OP) break;
            A.concops.push_back(m.concops[q]);
            A.concops.push_back(m.concops[q] == CACONC_OP ? m.concops[q + 1] + nconcs : m.concops[q + 1]);
        }
        G.row0.push_back(rows);
        G.col0.push_back(col0[i]);
        G.stab0.push_back(stab0[i]);
        G.conc0.push_back(nconcs);
        G.giv0.push_back(e.giv0);
        G.ra0.push_back((int32_t)A.ravals.size());
        G.funcs.push_back(m.funcs);
        G.nfuncs.push_back(m.nfuncs);
        G.ravals.push_back(m.ravals);
        G.nravals.push_back(m.nravals);
        G.nrows.push_back(m.ncompts);
        if (m.nravals > 0) A.ravals.insert(A.ravals.end(), m.ravals, m.ravals + m.nravals);
        G.member_of.insert(G.member_of.end(), (size_t)m.ncompts, i);
        rows += m.ncompts;
        nconcs += std::max(0, m.nconcs);
    }
    A.ops.push_back(LCOMPT_OP);
    if (!A.concops.empty()) A.concops.push_back(LCACONC_OP);
    hsh_model mm = m0;
    mm.ndiffs = mm.ncovals0 = mm.nbufvals = 0;
    mm.nshells = mm.nshrows = mm.nshfix = mm.dconc = mm.fbufval = mm.fdiffs = nullptr;
    mm.covals0 = mm.bufvals = nullptr;
    if (mx) {
        mm.xdivs = mx->xdivs;
        mm.xmin = mx->xmin;
        mm.invdx = mx->invdx;
    }
    mm.ydivs = ydivs;
    mm.ymin = my ? my->ymin : 0.0;
    mm.invdy = my ? my->invdy : 0.0;
    mm.ncompts = rows;
    mm.nchips = (int32_t)A.chip.size();
    mm.nops = (int32_t)A.ops.size();
    mm.ncols = ncols;
    mm.nstab = nstab;
    mm.nconcs = nconcs;
    mm.nconcops = (int32_t)A.concops.size();
    mm.nconcchip = (int32_t)A.concchip.size();
    mm.ngivals = (int32_t)A.giv.size();
    mm.nravals = (int32_t)A.ravals.size();
    mm.ops = A.ops.data();
    mm.chip = A.chip.data();
    mm.vm = A.vm.data();
    mm.tablist = ncols > 0 ? A.tab.data() : nullptr;
    mm.xvals = ncols > 0 ? A.xvals.data() : nullptr;
    mm.stablist = nstab > 0 ? A.stab.data() : nullptr;
    mm.ravals = A.ravals.empty() ? nullptr : A.ravals.data();
    mm.conc = nconcs > 0 ? A.conc.data() : nullptr;
    mm.concops = A.concops.empty() ? nullptr : A.concops.data();
    mm.concchip = A.concchip.empty() ? nullptr : A.concchip.data();
    mm.givals = A.giv.empty() ? nullptr : A.giv.data();
    // funcs[] are decoded per member; the merged model keeps the first's as a placeholder
    return create_impl(&mm, opt, out, std::move(ga), std::move(mem));
}

hsh_status hsh_group_member(hsh_handle* h, int32_t i, hsh_member* out)
{
    if (!h || !out) return fail(HSH_ERR_ARG, "hsh_group_member: null argument");
    if (i < 0 || i >= (int32_t)h->members.size()) return fail(HSH_ERR_ARG, "hsh_group_member: no such member");
    *out = h->members[i];
    return HSH_OK;
}

hsh_status hsh_refusals(hsh_handle* h, hsh_refusal* out, int32_t cap, int32_t* n)
{
    if (!h || !n) return fail(HSH_ERR_ARG, "hsh_refusals: null argument");
    const auto& r = h->plan.refusals;
    *n = (int32_t)r.size();
    if (out)
        for (int32_t i = 0; i < cap && i < (int32_t)r.size(); i++) out[i] = r[i];
    return HSH_OK;
}

hsh_status hsh_autotune(hsh_handle* h)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_autotune: null handle");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_autotune: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_autotune: handle refused the model");
    return autotune(h);
}

hsh_status hsh_step(hsh_handle* h, int32_t nsteps, int32_t mode, hsh_step_out* out)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_step: null handle");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_step: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_step: handle refused the model; use the CPU path");
    if (nsteps < 0) return fail(HSH_ERR_ARG, "hsh_step: negative step count");
    if (mode != HSH_PLAIN && mode != HSH_CHANNELS && mode != HSH_BATCHED)
        return fail(HSH_ERR_ARG, "hsh_step: unknown mode");
    if (mode == HSH_BATCHED && !h->tuned && !(h->opt.flags & HSH_F_NO_AUTOTUNE) && nsteps > 0)
        STCHK(autotune(h));
    hsh_step_out o{};
    STCHK(ev_flush(h));
    if (h->timing) HIPCHK(hipEventRecord(h->ev_a, h->stream));
    h->nev_used = 0;
    if (mode == HSH_PLAIN || mode == HSH_CHANNELS || h->plain_loop) {
        for (int s = 0; s < nsteps; s++) {
            STCHK(plain_step(h, mode == HSH_CHANNELS));
            o.launches += 2;
        }
        o.vm_ready = (mode != HSH_CHANNELS);
    } else {
        STCHK(run_batched(h, nsteps, &o.launches));
        o.vm_ready = 1;
    }
    if (h->timing) HIPCHK(hipEventRecord(h->ev_b, h->stream));
    h->total_valid = h->timing;
    h->host_step += nsteps;
    for (int s = 0; s < nsteps; s++) h->t_next += h->m.dt;
    if (out) *out = o;
    return HSH_OK;
}

hsh_status hsh_sync(hsh_handle* h)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_sync: null handle");
    if (h->refused || h->host_only) return HSH_OK;
    HIPCHK(hipStreamSynchronize(h->stream));
    return ev_check(h);
}

hsh_status hsh_get_state(hsh_handle* h, double* vm, double* chip, double* results, int32_t* ops)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_get_state: null handle");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_get_state: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_get_state: handle refused the model");
    STCHK(ev_flush(h));
    HIPCHK(hipStreamSynchronize(h->stream));
    STCHK(ev_check(h));
    if (vm || results) STCHK(download_vm_results(h, vm, results));
    if (chip || ops) {
        STCHK(download_chip_ops(h));
        if (chip) std::memcpy(chip, h->chip.data(), (size_t)h->m.nchips * sizeof(double));
        if (ops) std::memcpy(ops, h->ops.data(), h->ops.size() * sizeof(int32_t));
    }
    return HSH_OK;
}

hsh_status hsh_put_state(hsh_handle* h, const double* vm, const double* chip, const int32_t* ops)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_put_state: null handle");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_put_state: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_put_state: handle refused the model");
    STCHK(ev_flush(h));                 // posted events apply to the state being replaced
    HIPCHK(hipStreamSynchronize(h->stream));
    if (chip || ops) {
        // refresh the host copies with the device's mutable state first, so
        // passing only one of chip/ops keeps the other current
        STCHK(download_chip_ops(h));
    }
    if (chip) std::memcpy(h->chip.data(), chip, (size_t)h->m.nchips * sizeof(double));
    if (ops) std::memcpy(h->ops.data(), ops, h->ops.size() * sizeof(int32_t));
    if (chip || ops) return upload_state(h, vm);
    if (vm) {
        std::vector<double> V;
        pack_V(h, vm, V);
        HIPCHK(hipMemcpy(h->d_V, V.data(), V.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return HSH_OK;
}

hsh_status hsh_get_givals(hsh_handle* h, double* givals)
{
    if (!h || !givals) return fail(HSH_ERR_ARG, "hsh_get_givals: null argument");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_get_givals: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_get_givals: handle refused the model");
    if (!h->d_giv) return fail(HSH_ERR_STATE, "hsh_get_givals: the model was created without givals");
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(givals, h->d_giv, (size_t)h->m.ngivals * sizeof(double), hipMemcpyDeviceToHost));
    return HSH_OK;
}

hsh_status hsh_get_conc(hsh_handle* h, double* conc)
{
    if (!h || !conc) return fail(HSH_ERR_ARG, "hsh_get_conc: null argument");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_get_conc: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_get_conc: handle refused the model");
    HIPCHK(hipStreamSynchronize(h->stream));
    STCHK(download_chip_ops(h));
    std::memcpy(conc, h->chip.data() + h->m.nchips, (size_t)std::max(0, h->m.nconcs) * sizeof(double));
    return HSH_OK;
}

hsh_status hsh_put_conc(hsh_handle* h, const double* conc)
{
    if (!h || !conc) return fail(HSH_ERR_ARG, "hsh_put_conc: null argument");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_put_conc: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_put_conc: handle refused the model");
    HIPCHK(hipStreamSynchronize(h->stream));
    STCHK(download_chip_ops(h));
    std::memcpy(h->chip.data() + h->m.nchips, conc, (size_t)std::max(0, h->m.nconcs) * sizeof(double));
    return upload_state(h, nullptr);
}

}  // extern "C"

namespace {

// The device entries (position, set or add, value) of message inputs, in
// position order and the caller's order within a position, into buf
// (capacity cap; *cnt entries).  More than cap is HSH_ERR_LIMIT before
// anything changes.
hsh_status input_entries(hsh_handle* h, int32_t n, const hsh_input* in, hsh_handle::InTagged* buf, size_t cap,
                         size_t* cnt, const char* who)
{
    const int32_t nchips = h->m.nchips;
    size_t need = 0;
    for (int32_t i = 0; i < n; i++) {
        const hsh_input& e = in[i];
        if (e.op == HSH_IN_INJECT || e.op == HSH_IN_BASE || e.op == HSH_IN_THRESH) {
            if (e.index < 0 || e.index >= nchips)
                return fail(HSH_ERR_ARG, std::string(who) + ": chip index out of range");
            need += (size_t)(h->c2s_off[e.index + 1] - h->c2s_off[e.index]);
        } else if (e.op == HSH_IN_ACTIVATION) {
            if (e.index < 0 || e.index >= nchips || h->c2s_off[e.index + 1] - h->c2s_off[e.index] != 1)
                return fail(HSH_ERR_ARG, std::string(who) + ": activation index is not a synchan's X");
            if (e.stab < 0 || (size_t)e.stab + 3 >= h->stab_host.size())
                return fail(HSH_ERR_ARG, std::string(who) + ": stablist index out of range");
            need += 1;
        } else if (e.op == HSH_IN_FLUX) {
            if (e.index < 0 || e.index >= std::max(0, h->m.nconcs))
                return fail(HSH_ERR_ARG, std::string(who) + ": pool index out of range");
            if ((size_t)e.index < h->plan.flux_live.size() && h->plan.flux_live[e.index])
                return fail(HSH_ERR_ARG, std::string(who) + ": FLUX input into a shell whose FLUXCONC_OP reads a "
                                                            "flux no channel clears (not run on the device)");
        } else {
            return fail(HSH_ERR_ARG, std::string(who) + ": unknown input op " + std::to_string(e.op));
        }
    }
    if (need > cap)
        return fail(HSH_ERR_LIMIT, std::string(who) + ": " + std::to_string(need) + " device entries, more than " +
                                       std::to_string(cap));
    size_t k = 0;
    for (int32_t i = 0; i < n; i++) {
        const hsh_input& e = in[i];
        if (e.op == HSH_IN_INJECT || e.op == HSH_IN_BASE || e.op == HSH_IN_THRESH) {
            h->chip[e.index] = e.value;
            for (int64_t q = h->c2s_off[e.index]; q < h->c2s_off[e.index + 1]; q++, k++)
                buf[k] = hsh_handle::InTagged{InEntry{(int)h->c2s[q], 0, e.value}, (int32_t)k};
        } else if (e.op == HSH_IN_ACTIVATION) {
            if (!(e.value > 1.0e-10)) continue;          // VTINY (hines_defs.h:115)
            const double a = e.per_dt ? e.value * h->stab_host[e.stab + 3] * h->m.dt
                                      : e.value * h->stab_host[e.stab + 3];
            buf[k] = hsh_handle::InTagged{InEntry{(int)h->c2s[h->c2s_off[e.index]], 1, a}, (int32_t)k};
            k++;
        }
        // HSH_IN_FLUX: CONC_VAL_OP clears the pool's flux register at the
        // start of every program that feeds it (hines_chip.c:205-212), and
        // the device runs only pools such a program feeds (a shell whose
        // FLUXCONC_OP reads a flux nothing clears is refused above): the
        // value is overwritten before anything reads it
    }
    std::sort(buf, buf + k, [](const hsh_handle::InTagged& a, const hsh_handle::InTagged& b) {
        return a.e.pos != b.e.pos ? a.e.pos < b.e.pos : a.seq < b.seq;
    });
    *cnt = k;
    return HSH_OK;
}

}  // namespace

extern "C" {

hsh_status hsh_set_inputs(hsh_handle* h, int32_t n, const hsh_input* in)
{
    if (!h || (n > 0 && !in) || n < 0) return fail(HSH_ERR_ARG, "hsh_set_inputs: bad arguments");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_set_inputs: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_set_inputs: handle refused the model");
    size_t need = 0;
    for (int32_t i = 0; i < n; i++) {
        const hsh_input& e = in[i];
        const bool set = e.op == HSH_IN_INJECT || e.op == HSH_IN_BASE || e.op == HSH_IN_THRESH;
        if (set && e.index >= 0 && e.index < h->m.nchips) need += (size_t)(h->c2s_off[e.index + 1] - h->c2s_off[e.index]);
        else need += 1;
    }
    std::vector<hsh_handle::InTagged> tag(std::max<size_t>(1, need));
    size_t cnt = 0;
    STCHK(input_entries(h, n, in, tag.data(), tag.size(), &cnt, "hsh_set_inputs"));
    if (cnt == 0) return HSH_OK;
    std::vector<InEntry> e(cnt);
    for (size_t i = 0; i < cnt; i++) e[i] = tag[i].e;
    if ((int)cnt > h->inputs_cap) {
        HIPCHK(hipStreamSynchronize(h->stream));
        if (h->d_inputs) HIPCHK(hipFree(h->d_inputs));
        h->d_inputs = nullptr;
        h->inputs_cap = 0;
        int cap = 256;
        while (cap < (int)cnt) cap *= 2;
        HIPCHK(dalloc(&h->d_inputs, (size_t)cap));
        h->inputs_cap = cap;
    }
    HIPCHK(hipMemcpyAsync(h->d_inputs, e.data(), cnt * sizeof(InEntry), hipMemcpyHostToDevice, h->stream));
    HIPCHK(launch_inputs(h->d_S, h->d_inputs, (int)cnt, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return HSH_OK;
}

hsh_status hsh_set_inputs_static(hsh_handle* h, int32_t n, const hsh_input* in)
{
    if (!h || (n > 0 && !in) || n < 0) return fail(HSH_ERR_ARG, "hsh_set_inputs_static: bad arguments");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_set_inputs_static: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_set_inputs_static: handle refused the model");
    if (n > HSH_INPUTS_STATIC_MAX)
        return fail(HSH_ERR_LIMIT, "hsh_set_inputs_static: more than HSH_INPUTS_STATIC_MAX inputs");
    InEntry* scr = inputs_static_ptr(h->device);
    if (!scr) return fail(HSH_ERR_STATE, "hsh_set_inputs_static: static staging not set up on this device");
    size_t cnt = 0;
    STCHK(input_entries(h, n, in, h->in_tag.data(), h->in_tag.size(), &cnt, "hsh_set_inputs_static"));
    if (cnt == 0) return HSH_OK;
    for (size_t i = 0; i < cnt; i++) h->in_host[i] = h->in_tag[i].e;
    HIPCHK(hipMemcpyAsync(scr, h->in_host.data(), cnt * sizeof(InEntry), hipMemcpyHostToDevice, h->stream));
    HIPCHK(launch_inputs(h->d_S, scr, (int)cnt, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return HSH_OK;
}

hsh_status hsh_set_chip(hsh_handle* h, int32_t n, const int32_t* idx, const double* val)
{
    if (!h || (n > 0 && (!idx || !val))) return fail(HSH_ERR_ARG, "hsh_set_chip: null argument");
    std::vector<hsh_input> in((size_t)std::max(0, n));
    for (int32_t i = 0; i < n; i++) in[i] = hsh_input{HSH_IN_INJECT, idx[i], 0, 0, val[i]};
    hsh_status st = hsh_set_inputs(h, n, in.data());
    if (st != HSH_OK) g_err = "hsh_set_chip: " + g_err;
    return st;
}

hsh_status hsh_drain_spikes(hsh_handle* h, int32_t* opidx, int32_t* step, int32_t cap, int32_t* n)
{
    if (!h || !n) return fail(HSH_ERR_ARG, "hsh_drain_spikes: null argument");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_drain_spikes: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_drain_spikes: handle refused the model");
    HIPCHK(hipStreamSynchronize(h->stream));
    STCHK(ev_check(h));
    int32_t cnt = 0;
    HIPCHK(hipMemcpy(&cnt, h->d_nspk, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (cnt > SPIKE_CAP) return fail(HSH_ERR_LIMIT, "hsh_drain_spikes: device spike buffer overflowed; drain more often");
    if (cnt > cap) return fail(HSH_ERR_LIMIT, "hsh_drain_spikes: output capacity too small");
    std::vector<int32_t> buf(2 * (size_t)std::max(1, cnt));
    if (cnt > 0) HIPCHK(hipMemcpy(buf.data(), h->d_spk, 2 * (size_t)cnt * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<std::pair<int32_t, int32_t>> ev(cnt);
    for (int i = 0; i < cnt; i++) ev[i] = {buf[2 * i + 1], buf[2 * i]};
    std::sort(ev.begin(), ev.end());   // (step, ops index): the reference's op order
    for (int i = 0; i < cnt; i++) {
        if (step) step[i] = ev[i].first;
        if (opidx) opidx[i] = ev[i].second;
    }
    *n = cnt;
    HIPCHK(hipMemset(h->d_nspk, 0, sizeof(int32_t)));
    return HSH_OK;
}

hsh_status hsh_info_get(hsh_handle* h, hsh_info* info)
{
    if (!h || !info) return fail(HSH_ERR_ARG, "hsh_info_get: null argument");
    const Plan& P = h->plan;
    hsh_info o{};
    o.ntiles = (int32_t)P.tiles.size();
    o.nblocks = (int32_t)P.blocks.size();
    o.ntemplates = (int32_t)P.templates.size();
    o.ntrees = (int32_t)P.tree_lo.size();
    o.max_rows = P.max_rows;
    o.batch_k = std::min(h->batch_k, h->max_k);
    o.batch_waves = h->batch_waves;
    o.refused = h->refused ? 1 : 0;
    o.bytes_per_step = P.bytes_per_step;
    o.lane_fill = P.blocks.empty() ? 0.0 : (double)h->m.ncompts / ((double)P.blocks.size() * W);
    o.jit = h->jit.ok ? 1 : 0;
    o.npacks = (int32_t)P.packs.size();
    o.jit_compile_s = h->jit.compile_s;
    o.tile_max_rows = P.layout.tile_max_rows;
    o.pack_cap = P.layout.pack_cap;
    o.pack_waves = h->pack_waves;
    o.pack_stage = h->pack_stage + 1;
    o.max_tree_rows = P.max_tree_rows;
    o.autotuned = h->tuned ? 1 : 0;
    o.autotune_s = h->autotune_s;
    *info = o;
    return HSH_OK;
}

hsh_status hsh_set_timing(hsh_handle* h, int32_t on)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_set_timing: null handle");
    h->timing = on != 0;
    return HSH_OK;
}

hsh_status hsh_last_time(hsh_handle* h, int32_t kernel, double* ms)
{
    if (!h || !ms) return fail(HSH_ERR_ARG, "hsh_last_time: null argument");
    if (h->host_only || !h->total_valid)
        return fail(HSH_ERR_STATE, "hsh_last_time: timing was not enabled for the last hsh_step");
    HIPCHK(hipStreamSynchronize(h->stream));
    float t = 0.f;
    if (kernel == HSH_T_TOTAL) {
        HIPCHK(hipEventElapsedTime(&t, h->ev_a, h->ev_b));
        *ms = t;
        return HSH_OK;
    }
    double acc = 0;
    for (int i = 0; i + 2 < h->nev_used + 1 && i < h->nev_used; i += 3) {
        if (kernel == HSH_T_CHANNEL) HIPCHK(hipEventElapsedTime(&t, h->evs[i], h->evs[i + 1]));
        else HIPCHK(hipEventElapsedTime(&t, h->evs[i + 1], h->evs[i + 2]));
        acc += t;
    }
    *ms = acc;
    return HSH_OK;
}

hsh_status hsh_jit_source(const hsh_model* m, char* buf, int64_t cap, int64_t* len)
{
    if (!m || !len) return fail(HSH_ERR_ARG, "hsh_jit_source: null argument");
    Plan P;
    std::string err;
    if (!analyse(*m, P, err)) return fail(HSH_ERR_ARG, "hsh_jit_source: " + err);
    if (!P.refusals.empty()) return fail(HSH_REFUSED, "hsh_jit_source: model refused");
    const std::string src = jit_source(P, m->ncols);
    *len = (int64_t)src.size();
    if (buf && cap > 0) {
        const size_t n = std::min<size_t>((size_t)cap - 1, src.size());
        std::memcpy(buf, src.data(), n);
        buf[n] = '\0';
    }
    return HSH_OK;
}


hsh_status hsh_set_synapses(hsh_handle* h, int32_t n, const int32_t* src_op, const int32_t* dst_op,
                            const double* weight, const double* delay)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_set_synapses: null handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_set_synapses: handle refused the model");
    if (n < 0 || (n > 0 && (!src_op || !dst_op || !weight || !delay)))
        return fail(HSH_ERR_ARG, "hsh_set_synapses: bad arguments");
    const Plan& P = h->plan;
    std::map<int32_t, int32_t> src_id, dst_id;
    for (size_t i = 0; i < P.spike_ops.size(); i++) src_id[P.spike_ops[i]] = (int32_t)i;
    for (size_t i = 0; i < P.syn_ops.size(); i++) dst_id[P.syn_ops[i]] = (int32_t)i;
    const double dt = h->m.dt;
    std::vector<int32_t> csrc(n), cdst(n), cT(n);
    std::vector<double> cwdt(n);
    int min_T = 1 << 30;
    for (int32_t i = 0; i < n; i++) {
        auto s = src_id.find(src_op[i]);
        auto t = dst_id.find(dst_op[i]);
        if (s == src_id.end()) return fail(HSH_ERR_ARG, "hsh_set_synapses: src_op " + std::to_string(src_op[i]) + " is not a SPIKE_OP");
        if (t == dst_id.end()) return fail(HSH_ERR_ARG, "hsh_set_synapses: dst_op " + std::to_string(dst_op[i]) + " is not a SYN2_OP");
        // weight and delay are float fields of the synapse
        // (newconn_struct.h:158-159); an hsolve spike is sent at the current
        // time, so it hits (int)(delay/dt + 0.9) steps later
        // (newconn/synchan.c:193-194)
        const double dl = (double)(float)delay[i];
        if (!(dl >= 0.0) || dl / dt > 1e6) return fail(HSH_ERR_ARG, "hsh_set_synapses: delay out of range");
        const int T = (int)((dl - 0.0) / dt + 0.9);
        if (T < 1) return fail(HSH_ERR_ARG, "hsh_set_synapses: a connection shorter than one step");
        csrc[i] = s->second;
        cdst[i] = t->second;
        cT[i] = T;
        cwdt[i] = (double)(float)weight[i] / dt;
        min_T = std::min(min_T, T);
    }
    // An EVENT can hit as early as the next step (the head's re-timing
    // truncates, so its time field drifts upward and later events fire
    // early), so with lists on every step is its own launch after its spike
    // pass.
    (void)min_T;
    const int max_k = 1;
    if (h->host_only) {
        h->max_k = max_k;
        h->nconn = n;
        return HSH_OK;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    STCHK(ev_flush(h));
    STCHK(ev_rebuild(h, csrc, cdst, cT, cwdt));
    h->max_k = max_k;
    h->nconn = n;
    free_graphs(h);
    set_devmodel(h);
    return HSH_OK;
}

hsh_status hsh_post_events(hsh_handle* h, int32_t n, const int32_t* syn_op, const int32_t* steps,
                           const double* weight, int32_t after)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_post_events: null handle");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_post_events: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_post_events: handle refused the model");
    if (n < 0 || (n > 0 && (!syn_op || !steps || !weight)) || (after != 0 && after != 1))
        return fail(HSH_ERR_ARG, "hsh_post_events: bad arguments");
    if (after && h->host_step == 0)
        return fail(HSH_ERR_ARG, "hsh_post_events: no step has run to post after");
    const Plan& P = h->plan;
    std::map<int32_t, int32_t> dst_id;
    for (size_t i = 0; i < P.syn_ops.size(); i++) dst_id[P.syn_ops[i]] = (int32_t)i;
    std::vector<hsh_handle::Post> add;
    for (int32_t i = 0; i < n; i++) {
        auto t = dst_id.find(syn_op[i]);
        if (t == dst_id.end())
            return fail(HSH_ERR_ARG, "hsh_post_events: " + std::to_string(syn_op[i]) + " is not a SYN2_OP");
        if (steps[i] < 0) return fail(HSH_ERR_ARG, "hsh_post_events: negative step count");
        add.push_back(hsh_handle::Post{t->second, steps[i], after, (int32_t)(h->posts.size() + add.size()),
                                       (double)(float)weight[i] / h->m.dt});
    }
    if (n == 0) return HSH_OK;
    if (!h->ev_on) {
        // lists for every synchan, no connections
        HIPCHK(hipStreamSynchronize(h->stream));
        STCHK(ev_rebuild(h, {}, {}, {}, {}));
        h->max_k = 1;
        free_graphs(h);
        set_devmodel(h);
    }
    h->posts.insert(h->posts.end(), add.begin(), add.end());
    return HSH_OK;
}

hsh_status hsh_set_time(hsh_handle* h, double t)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_set_time: null handle");
    if (!(t == t)) return fail(HSH_ERR_ARG, "hsh_set_time: not a number");
    h->t_next = t;
    if (h->ev_on) {
        STCHK(ev_flush(h));
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpy(&h->d_evctl->tnext, &t, sizeof(double), hipMemcpyHostToDevice));
    }
    return HSH_OK;
}

hsh_status hsh_set_ravals(hsh_handle* h, const double* ravals)
{
    if (!h || !ravals) return fail(HSH_ERR_ARG, "hsh_set_ravals: null argument");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_set_ravals: handle refused the model");
    Plan& P = h->plan;
    for (int r = 0; r < h->m.ncompts; r++) {
        if (P.ra_fwd[r] >= 0) {
            P.ta[r] = ravals[P.ra_fwd[r]];
            P.tcf[r] = ravals[P.ra_fwd[r] + 1];
        }
        if (P.ra_bwd[r] >= 0) P.tcb[r] = ravals[P.ra_bwd[r]];
    }
    // the layout keeps its shape: only the coefficient arrays change
    const size_t npos = P.blocks.size() * W;
    std::vector<double> ta(npos), tcf(npos), tcb(npos);
    for (size_t q = 0; q < npos; q++) {
        const int g = P.pos_comp[q];
        ta[q] = P.ta[g];
        tcf[q] = P.tcf[g];
        tcb[q] = P.tcb[g];
    }
    for (const PackMeta& pm : P.packs) {
        const int base = pm.block0 * W;
        for (int j = pm.ke0; j < pm.ke0 + pm.nke; j++) {
            const int c = P.pos_comp[base + P.kidx[j]];
            P.kta[j] = P.ta[c];
            P.ktcf[j] = P.tcf[c];
        }
        for (int e = pm.be0; e < pm.be0 + pm.nbe; e++) {
            const int r = P.pos_comp[base + (P.bkp[e] & 0xffff)];
            P.btcb[e] = (P.bkp[e] >> 16) > 0 ? P.tcb[r] : 0.0;
        }
    }
    for (const Plan::FuncsUnit& u : P.funits)
        std::copy(ravals + u.qg, ravals + u.qg + u.nq, P.fravals.begin() + u.q0);
    if (h->host_only) return HSH_OK;
    HIPCHK(hipStreamSynchronize(h->stream));
    if (!P.funits.empty()) HIPCHK(dput(h->d_fwork, P.fravals));
    HIPCHK(dput(h->d_ta, ta));
    HIPCHK(dput(h->d_tcf, tcf));
    HIPCHK(dput(h->d_tcb, tcb));
    HIPCHK(dput(h->d_kta, P.kta));
    HIPCHK(dput(h->d_ktcf, P.ktcf));
    HIPCHK(dput(h->d_btcb, P.btcb));
    return HSH_OK;
}

hsh_status hsh_adj_forward(hsh_handle* h, int32_t nsteps, int32_t ckpt, int32_t nobs, const int32_t* obs,
                           double* trace)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_adj_forward: null handle");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_adj_forward: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_adj_forward: handle refused the model");
    if (nsteps <= 0 || ckpt <= 0 || nobs < 0 || (nobs > 0 && !obs))
        return fail(HSH_ERR_ARG, "hsh_adj_forward: bad step count, checkpoint interval or observed rows");
    {
        std::vector<char> seen(h->m.ncompts, 0);
        for (int j = 0; j < nobs; j++) {
            if (obs[j] < 0 || obs[j] >= h->m.ncompts) return fail(HSH_ERR_ARG, "hsh_adj_forward: observed row out of range");
            if (seen[obs[j]]++) return fail(HSH_ERR_ARG, "hsh_adj_forward: observed rows must be distinct");
        }
    }
    std::string why;
    if (!adj_supported(h->plan, why)) return fail(HSH_REFUSED, "hsh_adj_forward: the reverse program does not run " + why);
    STCHK(ensure_adj_layout(h));
    if (ckpt > nsteps) ckpt = nsteps;
    STCHK(adj_alloc(h, nsteps, ckpt, nobs));
    auto& A = h->adj;
    A.packs = !h->plan.packs.empty();
    std::vector<int32_t> pos(std::max(1, nobs));
    for (int j = 0; j < nobs; j++) pos[j] = h->plan.comp_pos[obs[j]];
    if (nobs > 0) HIPCHK(hipMemcpy(A.obs_pos, pos.data(), nobs * sizeof(int32_t), hipMemcpyHostToDevice));
    if (A.packs) STCHK(adj_pack_arrays(h, nobs, obs));
    const bool timing = h->timing;
    h->timing = false;
    HIPCHK(hipStreamSynchronize(h->stream));
    STCHK(adj_save(h, 0));
    if (A.packs) {
        // one batched launch per checkpoint segment, recording the trace
        DevModel dm = h->dm;
        dm.obs_idx = A.obs_idx;
        dm.nobs = nobs;
        for (int t0 = 0; t0 < nsteps; t0 += ckpt) {
            const int n = std::min(ckpt, nsteps - t0);
            dm.trace = nobs > 0 ? A.trace + (size_t)t0 * nobs : nullptr;
            HIPCHK(launch_forest(dm, &h->jit, n, h->pack_waves, h->pack_stage, pack_lds_bytes(h, h->pack_stage),
                                 h->stream));
            if (t0 + n < nsteps) STCHK(adj_save(h, (t0 + n) / ckpt));
        }
    } else {
        for (int t = 0; t < nsteps; t++) {
            STCHK(plain_step(h, false));
            HIPCHK(launch_gather(h->d_V, A.obs_pos, nobs, A.trace + (size_t)t * nobs, h->stream));
            if ((t + 1) % ckpt == 0 && t + 1 < nsteps) STCHK(adj_save(h, (t + 1) / ckpt));
        }
    }
    STCHK(adj_save(h, A.ncp - 1));
    h->timing = timing;
    h->host_step += nsteps;
    HIPCHK(hipStreamSynchronize(h->stream));
    if (trace && nobs > 0)
        HIPCHK(hipMemcpy(trace, A.trace, (size_t)nsteps * nobs * sizeof(double), hipMemcpyDeviceToHost));
    A.valid = true;
    return HSH_OK;
}

hsh_status hsh_adj_backward(hsh_handle* h, const double* dtrace, double* g_chip, double* g_ravals, double* g_vm0)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_adj_backward: null handle");
    auto& A = h->adj;
    if (h->host_only || h->refused || !A.valid)
        return fail(HSH_ERR_STATE, "hsh_adj_backward: no hsh_adj_forward run to differentiate");
    if (!dtrace && A.nobs > 0) return fail(HSH_ERR_ARG, "hsh_adj_backward: null dtrace");
    const Plan& P = h->plan;
    const size_t npos = A.npos;
    HIPCHK(hipStreamSynchronize(h->stream));
    if (A.nobs > 0)
        HIPCHK(hipMemcpy(A.dtrace, dtrace, (size_t)A.nsteps * A.nobs * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(A.LV, 0, npos * sizeof(double), h->stream));
    HIPCHK(hipMemsetAsync(A.LS, 0, A.nS * sizeof(double), h->stream));
    HIPCHK(hipMemsetAsync(A.GT, 0, 3 * npos * sizeof(double), h->stream));
    int32_t nspk0 = 0;
    HIPCHK(hipMemcpy(&nspk0, h->d_nspk, sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<double> giv0;
    STCHK(save_givals(h, giv0));
    const bool timing = h->timing;
    h->timing = false;
    AdjModel am{};
    am.d = h->dm;
    am.RB = A.RB;
    am.RD = A.RD;
    am.LV = A.LV;
    am.LS = A.LS;
    am.BB = A.BB;
    am.DB = A.DB;
    am.GT = A.GT;
    am.Dp = A.Dp;
    am.Bp = A.Bp;
    am.X = A.X;
    am.XB = A.XB;
    am.npos = (long long)npos;
    am.max_rows = P.max_rows;
    const int nseg = (A.nsteps + A.ckpt - 1) / A.ckpt;
    AdjPack ap{};
    ap.d = h->dm;
    ap.segV = A.segV;
    ap.segS = A.segS;
    ap.segB = A.segB;
    ap.segD = A.segD;
    ap.segX = A.segX;
    ap.npos = (long long)npos;
    ap.nS = (long long)A.nS;
    ap.dtrace = A.dtrace;
    ap.nobs = A.nobs;
    ap.obs_idx = A.obs_idx;
    ap.LV = A.LV;
    ap.LS = A.LS;
    ap.GT = A.GT;
    ap.ppar = A.ppar;
    ap.pta = A.pta;
    ap.ptcf = A.ptcf;
    ap.ptcb = A.ptcb;
    DevModel rec = h->dm;
    rec.recV = A.segV;
    rec.recS = A.segS;
    rec.recB = A.segB;
    rec.recD = A.segD;
    rec.recX = A.segX;
    rec.rec_npos = (long long)npos;
    rec.rec_nS = (long long)A.nS;
    if (A.packs) {
        // Per segment, last first: load its checkpoint and recompute it,
        // keeping every step's state and elimination (`stream`), then run
        // it backwards in one launch (adj_stream).  Segments alternate
        // between two buffers, so segment s-1 is recorded while segment s
        // runs backwards; the reverse launches run in segment order on their
        // own stream (the adjoints carry from one to the next).
        if (!h->adj_stream) {
            HIPCHK(hipStreamCreateWithFlags(&h->adj_stream, hipStreamNonBlocking));
            for (int b = 0; b < 2; b++) {
                HIPCHK(hipEventCreateWithFlags(&h->adj_rec[b], hipEventDisableTiming));
                HIPCHK(hipEventCreateWithFlags(&h->adj_rev[b], hipEventDisableTiming));
            }
        }
        HIPCHK(hipStreamSynchronize(h->stream));       // the cleared adjoints
        bool used[2] = {false, false};
        int last = 0;
        for (int sg = nseg - 1; sg >= 0; sg--) {
            const int t0 = sg * A.ckpt, t1 = std::min(A.nsteps, t0 + A.ckpt);
            const int b = sg & 1;
            const size_t ov = (size_t)b * A.ckpt * npos, os = (size_t)b * A.ckpt * A.nS;
            if (used[b]) HIPCHK(hipStreamWaitEvent(h->stream, h->adj_rev[b], 0));
            STCHK(adj_load(h, sg));
            DevModel rb = rec;
            rb.recV = A.segV + ov;
            rb.recS = A.segS + os;
            rb.recB = A.segB + ov;
            rb.recD = A.segD + ov;
            rb.recX = A.segX + ov;
            HIPCHK(launch_forest(rb, &h->jit, t1 - t0, h->pack_waves, h->pack_stage,
                                 pack_lds_bytes(h, h->pack_stage), h->stream));
            HIPCHK(hipEventRecord(h->adj_rec[b], h->stream));
            AdjPack ab = ap;
            ab.segV = A.segV + ov;
            ab.segS = A.segS + os;
            ab.segB = A.segB + ov;
            ab.segD = A.segD + ov;
            ab.segX = A.segX + ov;
            ab.t0 = t0;
            ab.n = t1 - t0;
            HIPCHK(hipStreamWaitEvent(h->adj_stream, h->adj_rec[b], 0));
            HIPCHK(launch_adj_forest(ab, &h->jit, P.max_pack_pos, h->adj_stream));
            HIPCHK(hipEventRecord(h->adj_rev[b], h->adj_stream));
            used[b] = true;
            last = b;
        }
        HIPCHK(hipStreamWaitEvent(h->stream, h->adj_rev[last], 0));
    }
    for (int sg = nseg - 1; sg >= 0 && !A.packs; sg--) {
        const int t0 = sg * A.ckpt, t1 = std::min(A.nsteps, t0 + A.ckpt);
        STCHK(adj_load(h, sg));
        for (int t = t0; t < t1; t++) {
            const size_t j = (size_t)(t - t0);
            HIPCHK(hipMemcpyAsync(A.segV + j * npos, h->d_V, npos * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
            HIPCHK(hipMemcpyAsync(A.segS + j * A.nS, h->d_S, (size_t)P.nS * sizeof(double), hipMemcpyDeviceToDevice,
                                  h->stream));
            STCHK(plain_step(h, false));
        }
        for (int t = t1 - 1; t >= t0; t--) {
            const size_t j = (size_t)(t - t0);
            HIPCHK(launch_inject(A.LV, A.obs_pos, A.nobs, A.dtrace + (size_t)t * A.nobs, h->stream));
            am.Vt = A.segV + j * npos;
            am.St = A.segS + j * A.nS;
            HIPCHK(launch_adj_rhs(am, h->stream));
            HIPCHK(launch_adj_elim(am, h->stream));
            HIPCHK(launch_adj_local(am, h->stream));
        }
    }
    STCHK(adj_load(h, A.ncp - 1));
    HIPCHK(hipMemcpyAsync(h->d_nspk, &nspk0, sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    STCHK(load_givals(h, giv0));
    h->timing = timing;
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<double> LV(npos), LS(A.nS), GT(3 * npos);
    HIPCHK(hipMemcpy(LV.data(), A.LV, npos * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(LS.data(), A.LS, A.nS * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(GT.data(), A.GT, 3 * npos * sizeof(double), hipMemcpyDeviceToHost));
    const int n = h->m.ncompts;
    if (g_chip) std::fill(g_chip, g_chip + h->m.nchips, 0.0);
    if (g_ravals) std::fill(g_ravals, g_ravals + h->m.nravals, 0.0);
    for (int g = 0; g < n; g++) {
        const int pos = P.comp_pos[g];
        const int b = pos / W, l = pos % W;
        if (g_vm0) g_vm0[g] = LV[pos];
        if (g_chip) {
            const BlockMeta& bm = P.blocks[b];
            const Template& t = P.templates[P.comp_tmpl[g]];
            for (int k = 0; k < t.nslots; k++) {
                const int64_t src = P.dsrc[P.comp_dsrc_off[g] + k];
                if (src >= 0) g_chip[src] += LS[(size_t)bm.dslot + (size_t)k * W + l];
            }
        }
        if (g_ravals) {
            if (P.ra_fwd[g] >= 0) {
                g_ravals[P.ra_fwd[g]] += GT[pos];
                g_ravals[P.ra_fwd[g] + 1] += GT[npos + pos];
            }
            if (P.ra_bwd[g] >= 0) g_ravals[P.ra_bwd[g]] += GT[2 * npos + pos];
        }
    }
    return HSH_OK;
}

void hsh_destroy(hsh_handle* h)
{
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_device(h);
    delete h;
}

}  // extern "C"
