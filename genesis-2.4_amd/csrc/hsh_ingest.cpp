// hsh_ingest.cpp -- host-side analysis of Hsolve arrays into the device layout.
//
// Three analysers:
//  * ops analyser: walks ops[] exactly as the reference interpreter does
//    (genesis/src/hines/hines_chip.c:157-600), resolves the registers that
//    the reference carries across compartment boundaries (the Ek register that
//    CHAN_OP inherits, hines_chip_init.c:1029-1049; the table-row cache,
//    hines_chip.c:178-190) and rewrites each compartment's program into device
//    micro-ops.  Compartments whose programs are identical share a template.
//    Anything the device interpreter cannot run is refused with a reason.
//  * tree analyser: decodes funcs[]/ravals[] (hines_solve.c:223-296, emitted
//    by hines_init.c:604-742, :1011-1172) into a parent-row array and the
//    three coefficients of each non-root row, and checks that every tree is a
//    contiguous row block whose kids are eliminated in ascending row order.
//  * layout (build_layout): groups neurons with the same size and the same
//    per-row templates into 64-lane tiles (lane = neuron), and packs larger
//    trees into workgroup-sized packs (lane = compartment) with a level
//    schedule for their elimination; every wave executes a uniform program.
#include <algorithm>
#include <cstring>
#include <map>

#include "hsh_internal.h"

namespace hsh {

namespace {

void refuse(Plan& p, int comp, int opcode, int opidx, int reason)
{
    p.refusals.push_back(hsh_refusal{comp, opcode, opidx, reason});
}

// Ek register state carried across compartments.
enum EkKind { EK_NONE, EK_CONST, EK_DYN };

}  // namespace

bool analyse(const hsh_model& m, Plan& P, std::string& err, const GroupSpec* g)
{
    const int n = m.ncompts;
    const int32_t* ops = m.ops;
    const double* chip = m.chip;
    if (n <= 0 || !ops || !chip || !m.vm || !m.funcs) { err = "empty or missing arrays"; return false; }
    if (m.calcmode != 1) {
        refuse(P, -1, 0, -1, HSH_R_CALCMODE);
    }

    // ------------------------------------------------------------------
    // ops analyser
    P.comp_tmpl.assign(n, -1);
    P.comp_dsrc_off.assign(n + 1, 0);
    P.comp_isrc_off.assign(n + 1, 0);
    std::map<std::vector<int32_t>, int> tmpl_index;
    std::vector<char> has_volt(n, 0);

    int64_t p = 0, c = 0;
    EkKind ek = EK_NONE;
    int64_t eksrc = -1;       // chip index of the CHAN_EK_OP value in the register
    int comp = -1;
    auto bad = [&](const char* what) {
        err = std::string("malformed ops[]: ") + what + " at ops index " + std::to_string(p);
        return false;
    };
    // Concentrations.  Slot sources beyond chip[]: nchips + k is conc[k],
    // nchips + nconcs + j is concchip[j], then givals, covals0 and bufvals.
    // Ca_concen pools (CACONC_OP k, hines_conc_chip.c:196-211) each take 3
    // concchip values and are updated at the end of the program that feeds
    // them.  The rest of concops[] -- the shells' results, pumps, pool
    // buffers and pools -- is cut into segments, one per run of ops on one
    // concentration (CONC_OP / FCONC_OP k ...), each run by the compartment
    // that owns k: the compartment of k's diffusion system (fdiffs), or for a
    // concpool the owner of the shell its Hill pump drains.
    const int64_t nchips = m.nchips, nconcs = std::max(0, m.nconcs);
    std::vector<int64_t> pool_chip(nconcs, -1);     // k -> first concchip value, -1: not a Ca_concen pool
    std::vector<int> pool_owner(nconcs, -1);       // k -> compartment whose program uses it
    std::vector<int> conc_owner(nconcs, -1);       // k -> compartment whose program updates it (shells, pools)
    const int nsys = m.ndiffs > 0 ? m.ndiffs : 0;
    std::vector<int> sys_owner(nsys, -1), conc_sys(nconcs, -1), conc_row(nconcs, -1);
    std::vector<int64_t> sys_cov(nsys, 0), sys_ncov(nsys, 0);
    std::vector<std::vector<int>> comp_sys(n);     // per compartment: the systems it owns
    const int64_t gsrc0 = nchips + nconcs + std::max(0, m.nconcchip);
    const bool giv_on = (m.givals != nullptr && m.ngivals > 0);
    const int64_t csrc0 = gsrc0 + (giv_on ? m.ngivals : 0);    // covals0[0]
    const int64_t bsrc0 = csrc0 + std::max(0, m.ncovals0);     // bufvals[0]
    bool diff_bad = false;
    if (nsys > 0) {
        if (g) {
            diff_bad = true;
        } else {
            for (int r = 0; r < n && !diff_bad; r++) {
                const int32_t a = m.fdiffs[r], b = m.fdiffs[r + 1];
                if (a < 0 || b < a || b > nsys) { diff_bad = true; break; }
                for (int32_t i = a; i < b; i++) {
                    if (sys_owner[i] >= 0) { diff_bad = true; break; }
                    sys_owner[i] = r;
                    comp_sys[r].push_back(i);
                }
            }
        }
        int64_t cur = 0;
        for (int i = 0; i < nsys && !diff_bad; i++) {
            const int ms = m.nshells[i], dw = m.nshrows[i], nd = m.nshfix[i];
            const int64_t c0 = m.dconc[i];
            if (sys_owner[i] < 0 || ms < 2 || dw < 1 || nd < 0 || nd > dw - 1 || (int64_t)ms * dw > 256 ||
                c0 < 0 || c0 + (int64_t)ms * dw > nconcs ||
                (dw > 1 && (m.fbufval[i] < 0 || m.fbufval[i] + 3LL * (dw - 1) > m.nbufvals))) {
                diff_bad = true;
                break;
            }
            int64_t cnt = 0;
            for (int s = 0; s < ms; s++) {
                const int km = (s == 0 || s == ms - 1) ? 1 : 2;
                cnt += nd + (int64_t)(dw - 1 - nd) * km + km;
            }
            sys_cov[i] = cur;
            sys_ncov[i] = cnt;
            cur += cnt;
            if (cur > m.ncovals0) { diff_bad = true; break; }
            for (int r = 0; r < ms * dw; r++) {
                if (conc_sys[c0 + r] >= 0) { diff_bad = true; break; }
                conc_sys[c0 + r] = i;
                conc_row[c0 + r] = r;
                conc_owner[c0 + r] = sys_owner[i];
            }
        }
        if (diff_bad) refuse(P, -1, 0, -1, HSH_R_DIFFUSION);
    }
    auto is_shell = [&](int64_t k) {
        return conc_sys[k] >= 0 && conc_row[k] % m.nshrows[conc_sys[k]] == m.nshrows[conc_sys[k]] - 1;
    };
    struct COp {
        int32_t code;
        int32_t x, hill;      // operand concentration (shell, buffer) / Hill coefficient / Vm row
        int64_t cc;           // first concchip value
    };
    struct CSeg {
        int32_t at;           // the concentration the ops apply to
        int owner;
        bool store;           // its sum is stored as at's coresult
        std::vector<COp> ops;
    };
    std::vector<CSeg> segs;
    std::vector<std::vector<int>> comp_segs(n);
    P.flux_live.assign(nconcs, 0);
    {
        int64_t q = 0, j = 0;
        int cur = -1;
        auto open_seg = [&](int32_t k) {
            segs.push_back(CSeg{k, -1, false, {}});
            cur = (int)segs.size() - 1;
        };
        auto need = [&](int64_t ops_n, int64_t chip_n) {
            return q + ops_n <= m.nconcops && j + chip_n <= m.nconcchip;
        };
        auto conc_ok = [&](int64_t k) { return k >= 0 && k < nconcs; };
        bool done = false;
        while (q < m.nconcops && !done) {
            const int32_t co = m.concops[q++];
            switch (co) {
            case CONC_OP:
            case FCONC_OP: {
                if (!need(1, 0) || !conc_ok(m.concops[q])) { err = "malformed concops[]"; return false; }
                if (co == CONC_OP && cur >= 0) segs[cur].store = true;
                open_seg(m.concops[q++]);
                break;
            }
            case LCONC_OP:
                if (cur >= 0) segs[cur].store = true;
                done = true;
                break;
            case LCACONC_OP:
                done = true;
                break;
            case CACONC_OP: {
                if (!need(1, 3) || !conc_ok(m.concops[q])) { err = "malformed concops[]"; return false; }
                pool_chip[m.concops[q++]] = j;
                j += 3;
                break;
            }
            case FLUXCONC_OP: case CTAUPUMP_OP: case TAUPUMP_OP: case MMPUMP_OP: case HILLPUMP_OP:
            case PHILLPUMP_OP: case BUFFER_OP: case POOL_OP: {
                const int nop = (co == TAUPUMP_OP || co == HILLPUMP_OP || co == BUFFER_OP) ? 1
                              : (co == PHILLPUMP_OP ? 2 : 0);
                const int nch = (co == TAUPUMP_OP) ? 5 : (co == PHILLPUMP_OP || co == BUFFER_OP) ? 3
                              : (co == POOL_OP ? 0 : 2);
                if (!need(nop, nch)) { err = "malformed concops[]"; return false; }
                if (cur < 0) open_seg(0);          // the reference's index starts at 0
                COp o{co, -1, 0, j};
                if (co == TAUPUMP_OP) o.x = m.concops[q];
                if (co == HILLPUMP_OP) o.hill = m.concops[q];
                if (co == PHILLPUMP_OP) { o.hill = m.concops[q]; o.x = m.concops[q + 1]; }
                if (co == BUFFER_OP) o.x = m.concops[q];
                if ((co == PHILLPUMP_OP || co == BUFFER_OP) && !conc_ok(o.x)) {
                    err = "malformed concops[]";
                    return false;
                }
                q += nop;
                j += nch;
                segs[cur].ops.push_back(o);
                break;
            }
            default:
                refuse(P, -1, co, (int)(q - 1), HSH_R_OPCODE);
                done = true;
                break;
            }
        }
        // owners: the diffusion system of the concentration, else (a concpool)
        // the owner of the shell its Hill pump drains
        for (size_t si = 0; si < segs.size() && !diff_bad; si++) {
            CSeg& sg = segs[si];
            if (sg.ops.empty() && (!sg.store || !is_shell(sg.at))) continue;
            int own = conc_owner[sg.at];
            if (own < 0)
                for (const COp& o : sg.ops)
                    if (o.code == PHILLPUMP_OP && conc_sys[o.x] >= 0) own = sys_owner[conc_sys[o.x]];
            bool ok = own >= 0;
            int32_t badop = sg.ops.empty() ? CONC_OP : sg.ops[0].code;
            for (const COp& o : sg.ops) {
                if (!ok) break;
                if ((o.code == CTAUPUMP_OP || o.code == TAUPUMP_OP || o.code == MMPUMP_OP ||
                     o.code == HILLPUMP_OP) && !is_shell(sg.at))
                    ok = false;                                                  // pumps act on shells
                if (o.code == TAUPUMP_OP && o.x != own) ok = false;              // Vm of another compartment
                if (o.code == PHILLPUMP_OP && (!is_shell(o.x) || conc_owner[o.x] != own)) ok = false;
                if (o.code == BUFFER_OP && (conc_sys[o.x] >= 0 || (conc_owner[o.x] >= 0 && conc_owner[o.x] != own)))
                    ok = false;
                if (!ok) badop = o.code;
            }
            if (ok && conc_sys[sg.at] < 0 && conc_owner[sg.at] >= 0 && conc_owner[sg.at] != own) ok = false;
            if (!ok) {
                refuse(P, own, badop, -1, HSH_R_DIFFUSION);
                continue;
            }
            sg.owner = own;
            // a concpool and its buffers are updated by this owner only
            bool upd = false;
            for (const COp& o : sg.ops) {
                if (o.code == POOL_OP) upd = true;
                if (o.code == BUFFER_OP) conc_owner[o.x] = own;
            }
            if (upd && conc_sys[sg.at] < 0) conc_owner[sg.at] = own;
            comp_segs[own].push_back((int)si);
        }
    }
    // shared synthetic initial values of the solve's per-step slots
    int64_t synth_one = 0, synth_zero = 0;
    if (nsys > 0 && !diff_bad) {
        P.synth.push_back(1.0);
        synth_one = -(int64_t)P.synth.size();
        P.synth.push_back(0.0);
        synth_zero = -(int64_t)P.synth.size();
    }
    int creg = -1;            // conc register (CONC_VAL_OP), carried across compartments
    // givals: written in op order from givals[1] (hines_chip.c:114); value-space
    // sources of its entries follow concchip[]
    P.giv = giv_on;
    P.chanmode5 = (m.chanmode == 5);
    int64_t gcur = 1;
    // member offsets of a group (all zero for one solver)
    int32_t ocol = 0, ostab = 0, oconc = 0;
    auto member_start = [&](int r) {
        return g && (r == 0 || g->member_of[r] != g->member_of[r - 1]);
    };
    for (;;) {
        if (p >= m.nops) return bad("missing LCOMPT_OP");
        int32_t head = ops[p++];
        if (head == LCOMPT_OP) break;
        if (head != COMPT_OP && head != FCOMPT_OP) return bad("expected COMPT_OP");
        comp++;
        if (comp >= n) return bad("more compartments than ncompts");
        if (member_start(comp)) {
            // a new solver: its chip update starts with fresh registers
            // (hines_chip.c:106-150)
            const int mi = g->member_of[comp];
            ek = EK_NONE;
            eksrc = -1;
            creg = -1;
            gcur = g->giv0[mi] + 1;
            ocol = g->col0[mi];
            ostab = g->stab0[mi];
            oconc = g->conc0[mi];
        }
        std::vector<int32_t> code;
        std::vector<int64_t> dsrc;
        std::vector<int64_t> isrc;
        bool refused = false;
        int ngates = 0;
        bool ek_intra = false;    // Ek register set in this compartment
        bool creg_intra = false;  // conc register set in this compartment
        bool crow_intra = false;  // conc table row set in this compartment
        bool tabr_intra = false;  // table register (last lookup position) set in this compartment
        bool ipol_intra = false;  // first-concentration fraction (NEWCONC2) set in this compartment
        bool ghk_intra = false;   // GHK slope register set in this compartment
        bool tabghk = false;      // DO_TABGHK_OP: the next table lookup sets Ek and the slope
        // slots after tbyc/diagterm: pool concentrations, outside concentrations,
        // pool constants; code positions that hold such a slot index
        std::vector<int64_t> extra;
        std::vector<std::pair<size_t, int>> fix;
        std::vector<int> pools;   // local pool number -> k
        auto extra_slot = [&](int64_t src) {
            for (size_t i = 0; i < extra.size(); i++)
                if (extra[i] == src) return (int)i;
            extra.push_back(src);
            return (int)extra.size() - 1;
        };
        auto pool_of = [&](int k) {
            for (size_t i = 0; i < pools.size(); i++)
                if (pools[i] == k) return (int)i;
            pools.push_back(k);
            return (int)pools.size() - 1;
        };
        auto push_extra = [&](int64_t src) {
            fix.emplace_back(code.size(), extra_slot(src));
            code.push_back(-1);
        };
        auto raw_slot = [&](int64_t src) {
            extra.push_back(src);
            return (int)extra.size() - 1;
        };
        auto push_fixed = [&](int idx) {
            fix.emplace_back(code.size(), idx);
            code.push_back(-1);
        };
        // true when the conductance ops from ops index q set Ek from a table
        // (power-0 IPOL) before anything reads it
        auto ek_dead = [&](int64_t q) {
            while (q + 2 < m.nops) {
                const int32_t k = ops[q];
                if (k != IPOL1V_OP && k != IPOL1C_OP && k != IPOL2V_OP && k != IPOL2C_OP) return false;
                if (ops[q + 2] == 0) return true;
                q += 3;
            }
            return false;
        };
        EkKind ek_in = ek;        // live-in Ek
        (void)0;
        if (c + 2 > m.nchips) return bad("chip[] too short");
        dsrc.push_back(c); dsrc.push_back(c + 1);   // Em/Rm, inject
        c += 2;
        if (P.giv) {
            // leak current from Rm (givals[g]) into givals[g+1] (hines_chip.c:164-165)
            if (gcur + 2 > m.ngivals) return bad("givals[] too short");
            code.push_back(U_GIV0);
            push_extra(gsrc0 + gcur);
            code.push_back((int32_t)isrc.size());
            isrc.push_back(-(1 + gcur));
            gcur += 2;
        }
        auto need_chip = [&](int k) { return c + k <= m.nchips; };
        // The diffusion systems this compartment owns: their slots first, each
        // group contiguous -- concentrations (state), covals0 values, buffer
        // rates, then per step codiag and coresult per shell and the
        // elimination scratch (values within the step only).
        struct SysSlots {
            int sys, conc, cov, buf, dg, cr, x, r;
        };
        std::vector<SysSlots> sys_slots;
        int nscratch = 0;
        if (!diff_bad && !g)
            for (int i : comp_sys[comp]) {
                const int ms = m.nshells[i], dw = m.nshrows[i], N = ms * dw;
                SysSlots ss;
                ss.sys = i;
                ss.conc = (int)extra.size();
                for (int r = 0; r < N; r++) raw_slot(nchips + m.dconc[i] + r);
                ss.cov = (int)extra.size();
                for (int64_t l = 0; l < sys_ncov[i]; l++) raw_slot(csrc0 + sys_cov[i] + l);
                ss.buf = (int)extra.size();
                if (dw > 1)
                    for (int q = 0; q < 3 * (dw - 1); q++) raw_slot(bsrc0 + m.fbufval[i] + q);
                ss.dg = (int)extra.size();
                for (int q = 0; q < ms; q++) raw_slot(synth_one);
                ss.cr = (int)extra.size();
                for (int q = 0; q < ms; q++) raw_slot(synth_zero);
                ss.x = (int)extra.size();
                const int nx = dw == 1 ? ms : N * dw;
                for (int q = 0; q < nx; q++) raw_slot(synth_zero);
                ss.r = (int)extra.size();
                for (int q = 0; q < N; q++) raw_slot(synth_zero);
                nscratch += 2 * ms + nx + N;
                ngates += N;
                sys_slots.push_back(ss);
            }
        // the codiag / coresult slot of a shell of an owned system, -1
        auto shell_slot = [&](int64_t k, bool diag) {
            if (k < 0 || k >= nconcs || !is_shell(k) || conc_owner[k] != comp) return -1;
            for (const SysSlots& ss : sys_slots)
                if (ss.sys == conc_sys[k]) return (diag ? ss.dg : ss.cr) + conc_row[k] / m.nshrows[ss.sys];
            return -1;
        };
        while (p < m.nops && ops[p] > LCOMPT_OP) {
            int64_t opidx = p;
            int32_t op = ops[p++];
            bool to_cond = false;
            switch (op) {
            case NEWVOLT_OP:
                code.push_back(U_NEWVOLT);
                has_volt[comp] = 1;
                break;
            case SPIKE_OP:
                if (!need_chip(1) || p + 2 > m.nops) return bad("SPIKE_OP operands");
                code.push_back(U_SPIKE);
                code.push_back((int32_t)isrc.size());
                isrc.push_back(p);                   // counter (mutable)
                isrc.push_back(p + 1);               // reload
                isrc.push_back(-(1 + opidx));        // ops index, literal
                isrc.push_back(-(1 + (int64_t)P.spike_ops.size()));   // spike source id, literal
                P.spike_ops.push_back((int32_t)opidx);
                P.spike_comp.push_back(comp);
                P.spike_thresh.push_back(c);
                dsrc.push_back(c);                   // threshold
                c += 1;
                p += 2;
                break;
            case CONC_VAL_OP: {
                if (p >= m.nops) return bad("CONC_VAL_OP operand");
                const int32_t k = ops[p++] + oconc;
                if (k < 0 || k >= nconcs) return bad("CONC_VAL_OP pool index out of range");
                creg = k;
                creg_intra = true;
                if (pool_owner[k] < 0) pool_owner[k] = comp;
                if ((pool_owner[k] != comp && pool_chip[k] >= 0) || (conc_owner[k] >= 0 && conc_owner[k] != comp)) {
                    // a pool several compartments feed: the device keeps a
                    // pool's flux in the program of the compartment that owns it
                    refuse(P, comp, op, (int)opidx, HSH_R_POOL);
                    refused = true;
                    break;
                }
                const int pn = pool_of(k);
                if (pn >= MAXPOOL) { refuse(P, comp, op, (int)opidx, HSH_R_POOL); refused = true; break; }
                code.push_back(U_CONCVAL);
                code.push_back(pn);
                break;
            }
            case NEWCONC1_OP:
                if (!creg_intra) { refuse(P, comp, op, (int)opidx, HSH_R_REGISTER_LIVEIN); refused = true; break; }
                if (m.ydivs <= 0) return bad("NEWCONC1_OP without a concentration table");
                code.push_back(U_NEWCONC1);
                push_extra(nchips + creg);
                crow_intra = true;
                break;
            case DOMAIN_OP:
                if (!need_chip(1)) return bad("chip[] too short");
                if (m.ydivs <= 0) return bad("DOMAIN_OP without a concentration table");
                code.push_back(U_DOMAIN);
                dsrc.push_back(c);
                c += 1;
                crow_intra = true;
                break;
            case NEWCONC2_OP: {
                // the C1 row becomes the table register, then the row of
                // conc[k] (hines_chip.c:224-232)
                if (p >= m.nops) return bad("NEWCONC2_OP operand");
                const int32_t k = ops[p++] + oconc;
                if (k < 0 || k >= nconcs) return bad("NEWCONC2_OP pool index out of range");
                if (m.ydivs <= 0) return bad("NEWCONC2_OP without a concentration table");
                if ((pool_chip[k] >= 0 && pool_owner[k] != comp) || (conc_owner[k] >= 0 && conc_owner[k] != comp)) {
                    refuse(P, comp, op, (int)opidx, HSH_R_POOL); refused = true;
                    break;
                }
                code.push_back(U_NEWCONC2);
                push_extra(nchips + k);
                tabr_intra = ipol_intra = crow_intra;
                crow_intra = true;
                break;
            }
            case ELMMPUMP_OP:
                // electrogenic pump on the concentration of the last CONC_VAL_OP
                // (hines_chip.c:278-302): a current, and a diagonal term for
                // that shell's solve
                if (!need_chip(3)) return bad("chip[] too short");
                if (!creg_intra) {
                    refuse(P, comp, op, (int)opidx, HSH_R_REGISTER_LIVEIN); refused = true;
                } else if (conc_owner[creg] >= 0 && conc_owner[creg] != comp) {
                    refuse(P, comp, op, (int)opidx, HSH_R_POOL); refused = true;
                } else {
                    code.push_back(U_ELMMPUMP);
                    push_extra(nchips + creg);
                    const int dgs = shell_slot(creg, true);
                    if (dgs >= 0) push_fixed(dgs); else code.push_back(-1);
                    for (int i = 0; i < 3; i++) dsrc.push_back(c + i);
                }
                c += 3;
                if (P.giv) gcur += 1;
                break;
            case FLUX_OP:
            case SCALE_FLUX_OP:
                if (!creg_intra) { refuse(P, comp, op, (int)opidx, HSH_R_REGISTER_LIVEIN); refused = true;
                                   if (op == SCALE_FLUX_OP) c += 1; break; }
                code.push_back(op == FLUX_OP ? U_FLUX : U_SCALEFLUX);
                code.push_back(pool_of(creg));
                if (op == SCALE_FLUX_OP) {
                    if (!need_chip(1)) return bad("chip[] too short");
                    dsrc.push_back(c);
                    c += 1;
                }
                break;
            case STORE_ALLI_OP: case STORE_ALLC_OP:
                p += 1; refuse(P, comp, op, (int)opidx, HSH_R_OPCODE); refused = true; break;
            case GHK_OP:
            case NERNST_OP: {
                const int nc = (op == GHK_OP) ? 3 : 2;
                if (!need_chip(nc)) return bad("chip[] too short");
                if (creg_intra && creg + 1 >= nconcs) return bad("no outside concentration after the pool");
                if (!creg_intra) {
                    refuse(P, comp, op, (int)opidx, HSH_R_REGISTER_LIVEIN); refused = true;
                } else if (pool_chip[creg + 1] >= 0 || (conc_owner[creg + 1] >= 0 && conc_owner[creg + 1] != comp)) {
                    // the outside concentration is a pool another program updates
                    refuse(P, comp, op, (int)opidx, HSH_R_POOL); refused = true;
                } else {
                    code.push_back(op == GHK_OP ? U_GHK : U_NERNST);
                    if (op == GHK_OP) ghk_intra = true;
                    push_extra(nchips + creg);
                    push_extra(nchips + creg + 1);
                    for (int i = 0; i < nc; i++) dsrc.push_back(c + i);
                }
                c += nc;
                ek = EK_DYN; ek_intra = true;
                to_cond = true;
                break;
            }
            case SAME_GHK_OP:
                // gmax times the slope of the last GHK_OP / table GHK (hines_chip.c:358-365)
                if (!need_chip(1)) return bad("chip[] too short");
                if (!ghk_intra) {
                    refuse(P, comp, op, (int)opidx, HSH_R_REGISTER_LIVEIN); refused = true;
                } else {
                    code.push_back(U_SAMEGHK);
                    dsrc.push_back(c);
                }
                c += 1;
                to_cond = true;
                break;
            case DO_TABGHK_OP:
                tabghk = true;
                to_cond = true;
                break;
            case CHAN_EK_OP:
                if (!need_chip(2)) return bad("chip[] too short");
                code.push_back(U_CHAN_EK);
                dsrc.push_back(c); dsrc.push_back(c + 1);
                ek = EK_CONST; eksrc = c; ek_intra = true;
                c += 2;
                to_cond = true;
                break;
            case CHAN_OP:
                if (!need_chip(1)) return bad("chip[] too short");
                if ((ek_intra && ek == EK_DYN) || ek_dead(p)) {
                    // Ek computed in this compartment, or overwritten by a
                    // table current before the current is added
                    code.push_back(U_CHAN);
                    dsrc.push_back(c);
                } else if (!ek_intra && ek_in == EK_DYN) {
                    refuse(P, comp, op, (int)opidx, HSH_R_EK_LIVEIN);
                    refused = true;
                } else {
                    // The register holds a constant: the chip value of an
                    // earlier CHAN_EK_OP (this compartment or a previous one),
                    // or its initial 0.0.  Materialise it as an explicit slot
                    // that reads the same chip value.
                    code.push_back(U_CHAN_EK);
                    if (eksrc >= 0) {
                        dsrc.push_back(eksrc);
                    } else {
                        P.synth.push_back(0.0);
                        dsrc.push_back(-(int64_t)P.synth.size());
                    }
                    dsrc.push_back(c);
                    ek = EK_CONST; ek_intra = true;
                }
                c += 1;
                to_cond = true;
                break;
            default:
                refuse(P, comp, op, (int)opidx, HSH_R_OPCODE);
                err = "undefined chip operation " + std::to_string(op) + " at ops index " +
                      std::to_string(opidx);
                return false;
            }
            if (!to_cond) continue;
            // conductance ops until the current is added (hines_chip.c:404-590)
            for (;;) {
                if (p >= m.nops) return bad("unterminated channel");
                int64_t kidx = p;
                int32_t k = ops[p++];
                if (k == IPOL1V_OP || k == IPOL1C_OP || k == IPOL2V_OP || k == IPOL2C_OP) {
                    if (p + 2 > m.nops) return bad("IPOL operands");
                    int32_t col = ops[p] + ocol, pw = ops[p + 1];
                    p += 2;
                    if (pw != 0 && !need_chip(1)) return bad("chip[] too short");
                    if (pw < -6 || pw > 6) return bad("gate power out of range");
                    const bool twod = (k == IPOL2V_OP || k == IPOL2C_OP);
                    if (col < 0 || col + (twod ? 3 : 1) >= m.ncols || !m.tablist)
                        return bad("IPOL table column out of range");
                    // which table: 0 voltage row, 1 concentration row, 2 voltage x
                    // concentration, 3 two concentrations through the table register
                    const int which = k == IPOL1V_OP ? 0 : k == IPOL1C_OP ? 1 : k == IPOL2V_OP ? 2 : 3;
                    bool live = which == 0 ? has_volt[comp] : crow_intra;
                    if (which == 2) live = live && has_volt[comp];
                    if (which == 3) live = live && ipol_intra && tabr_intra;
                    if (!live) {
                        refuse(P, comp, k, (int)kidx, HSH_R_TABLE_LIVEIN); refused = true;
                    }
                    tabr_intra = true;
                    if (tabghk) {
                        // table GHK (hines_chip.c:473-477): Ek and the slope
                        // from the table, no current yet
                        tabghk = false;
                        if (pw != 0) {
                            refuse(P, comp, DO_TABGHK_OP, (int)kidx, HSH_R_OPCODE); refused = true;
                            c += 1;
                            continue;
                        }
                        code.push_back(U_TABGHK);
                        code.push_back(col);
                        code.push_back(which);
                        ghk_intra = true;
                        ek = EK_DYN; ek_intra = true;
                        break;
                    }
                    if (pw == 0) {
                        // table current (hines_chip.c:472-481): Ek from the
                        // table, Gk scaled by it, the current added
                        code.push_back(U_TABCURR);
                        code.push_back(col);
                        code.push_back(which);
                        ek = EK_DYN; ek_intra = true;
                        if (P.giv) gcur += 3;
                        break;
                    }
                    if (which <= 1) {
                        code.push_back(which == 1 ? U_IPOL1C : U_IPOL1V);
                    } else {
                        code.push_back(U_IPOLG);
                        code.push_back(which);
                    }
                    code.push_back(col);
                    code.push_back(pw);
                    dsrc.push_back(c);
                    c += 1;
                    ngates++;
                    continue;
                } else if (k == SYN2_OP || k == SYN3_OP) {
                    // [SYN2_OP, stablist index, event counter, child] (hines_chip.c:511-550)
                    if (p + 3 > m.nops) return bad("SYN2_OP operands");
                    const int32_t st = ops[p] + ostab;
                    const int64_t cntidx = p + 1;
                    const bool mg = (p + 3 < m.nops && ops[p + 3] == MGBLOCK_OP);
                    if (k == SYN3_OP) {
                        // SYN3 draws urandom() in the reference's serial op order
                        c += 1;
                        p += 3;
                        c += 2;
                        if (mg) { p += 1; c += 3; }
                        refuse(P, comp, k, (int)kidx, HSH_R_OPCODE); refused = true;
                        break;
                    }
                    if (!need_chip(mg ? 5 : 2)) return bad("chip[] too short");
                    if (st < 0 || (int64_t)st + 3 >= 4 * (int64_t)m.nstab || !m.stablist)
                        return bad("SYN2_OP stablist index out of range");
                    if (ops[cntidx] >= 0) {
                        refuse(P, comp, k, (int)kidx, HSH_R_SYN_PENDING); refused = true;
                    }
                    code.push_back(mg ? U_SYN2MG : U_SYN2);
                    code.push_back(st);
                    code.push_back((int32_t)isrc.size());
                    isrc.push_back(cntidx);                              // counter (mutable)
                    isrc.push_back(-(1 + (int64_t)P.syn_ops.size()));    // synchan id, literal
                    P.syn_ops.push_back((int32_t)kidx);
                    dsrc.push_back(c);          // X
                    dsrc.push_back(c + 1);      // Y
                    c += 2;
                    p += 3;
                    ngates += 2;
                    if (mg) {                   // KMg_A, 1/KMg_B, [Mg] (hines_chip_init.c:1180-1187)
                        for (int i = 0; i < 3; i++) dsrc.push_back(c + i);
                        c += 3;
                        p += 1;
                    }
                    if (P.giv) gcur += mg ? 6 : 3;   // the unblocked values first (hines_chip.c:533-542)
                    break;
                } else if (k == ADD_CURR_OP) {
                    code.push_back(U_ADDCURR);
                    if (P.giv) gcur += 3;
                    break;
                } else {
                    err = "undefined chip operation #2 " + std::to_string(k) + " at ops index " +
                          std::to_string(kidx);
                    return false;
                }
            }
            if (tabghk) {
                // DO_TABGHK_OP not followed by a power-0 table lookup
                refuse(P, comp, DO_TABGHK_OP, -1, HSH_R_OPCODE); refused = true;
                tabghk = false;
            }
        }
        if (c + 2 > m.nchips) return bad("chip[] too short at compartment end");
        // the Ca_concen update of the pools this compartment owns, after all
        // of its ops have read and fed them (the reference updates pools after
        // the whole chip update, hines.c:222-270; a pool owned by one
        // compartment sees no difference)
        // its concentration program (hines_conc_chip.c) and diffusion solves
        // (hines_conc.c), after its channel ops as the reference's passes come
        // after the whole chip update: nothing another compartment reads
        if (!refused) {
            std::vector<int64_t> upd;           // pool / buffer concentrations it updates
            auto state_conc = [&](int64_t k) {
                const int sl = extra_slot(nchips + k);
                if (conc_sys[k] < 0 && std::find(upd.begin(), upd.end(), k) == upd.end()) upd.push_back(k);
                return sl;
            };
            for (int si : comp_segs[comp]) {
                const CSeg& sg = segs[si];
                code.push_back(U_CZERO);
                const int dgx = shell_slot(sg.at, true);
                auto vals = [&](int64_t j0, int cnt) {
                    const int first = (int)extra.size();
                    for (int q = 0; q < cnt; q++) raw_slot(nchips + nconcs + j0 + q);
                    return first;
                };
                for (const COp& o : sg.ops) {
                    switch (o.code) {
                    case FLUXCONC_OP: {
                        int pn = -1;
                        for (size_t i = 0; i < pools.size(); i++)
                            if (pools[i] == sg.at) pn = (int)i;
                        code.push_back(U_FLUXCONC);
                        push_fixed(vals(o.cc, 2));
                        code.push_back(pn);
                        if (pn < 0) P.flux_live[sg.at] = 1;
                        break;
                    }
                    case CTAUPUMP_OP:
                        code.push_back(U_CTAUPUMP);
                        push_fixed(vals(o.cc, 2));
                        push_fixed(dgx);
                        break;
                    case TAUPUMP_OP:
                        code.push_back(U_TAUPUMP);
                        push_fixed(vals(o.cc, 5));
                        push_fixed(dgx);
                        break;
                    case MMPUMP_OP:
                        code.push_back(U_MMPUMP);
                        push_fixed(vals(o.cc, 2));
                        push_extra(nchips + sg.at);
                        push_fixed(dgx);
                        break;
                    case HILLPUMP_OP:
                        code.push_back(U_HILLPUMP);
                        push_fixed(vals(o.cc, 2));
                        push_extra(nchips + sg.at);
                        push_fixed(dgx);
                        code.push_back(o.hill);
                        break;
                    case PHILLPUMP_OP:
                        code.push_back(U_PHILLPUMP);
                        push_fixed(vals(o.cc, 3));
                        push_extra(nchips + o.x);
                        push_fixed(shell_slot(o.x, true));
                        code.push_back(o.hill);
                        break;
                    case BUFFER_OP:
                        code.push_back(U_BUFFER);
                        push_fixed(vals(o.cc, 3));
                        push_fixed(state_conc(o.x));
                        push_extra(nchips + sg.at);
                        break;
                    case POOL_OP:
                        code.push_back(U_POOL);
                        push_fixed(state_conc(sg.at));
                        break;
                    default:
                        break;
                    }
                }
                const int crs = shell_slot(sg.at, false);
                if (sg.store && crs >= 0) {
                    code.push_back(U_CSTORE);
                    push_fixed(crs);
                }
            }
            for (const SysSlots& ss : sys_slots) {
                code.push_back(U_DIFF);
                code.push_back(m.nshells[ss.sys]);
                code.push_back(m.nshrows[ss.sys]);
                code.push_back(m.nshfix[ss.sys]);
                for (int b : {ss.conc, ss.cov, ss.buf, ss.dg, ss.cr, ss.x, ss.r}) push_fixed(b);
            }
            ngates += (int)upd.size();
        }
        for (size_t pn = 0; pn < pools.size() && !refused; pn++) {
            const int k = pools[pn];
            if (pool_chip[k] < 0) continue;
            code.push_back(U_POOLCA);
            code.push_back((int32_t)pn);
            push_extra(nchips + k);
            push_extra(nchips + nconcs + pool_chip[k]);
            extra_slot(nchips + nconcs + pool_chip[k] + 1);
            extra_slot(nchips + nconcs + pool_chip[k] + 2);
        }
        code.push_back(U_END);
        dsrc.push_back(c); dsrc.push_back(c + 1);   // tbyc, diagterm
        c += 2;
        if (P.giv) gcur += 1;                       // total membrane current
        if (refused) continue;
        {
            // the three pool constants must be consecutive slots
            const int base = (int)dsrc.size();
            for (auto& f : fix) code[f.first] = base + f.second;
            for (int64_t src : extra) dsrc.push_back(src);
            for (size_t pn = 0; pn < pools.size(); pn++)
                if (pool_chip[pools[pn]] >= 0) ngates++;      // the pool concentration is state
        }
        Template t;
        t.code = code;
        t.nslots = (int)dsrc.size();
        t.nislots = (int)isrc.size();
        t.ngates = ngates;
        t.nscratch = nscratch;
        std::vector<int32_t> key = code;
        key.push_back(t.nslots);
        key.push_back(t.nislots);
        auto it = tmpl_index.find(key);
        int ti;
        if (it == tmpl_index.end()) {
            ti = (int)P.templates.size();
            tmpl_index[key] = ti;
            P.templates.push_back(t);
        } else {
            ti = it->second;
        }
        P.comp_tmpl[comp] = ti;
        P.comp_dsrc_off[comp] = (int64_t)P.dsrc.size();
        P.comp_isrc_off[comp] = (int64_t)P.isrc.size();
        P.dsrc.insert(P.dsrc.end(), dsrc.begin(), dsrc.end());
        P.isrc.insert(P.isrc.end(), isrc.begin(), isrc.end());
    }
    if (comp != n - 1) { err = "ops[] holds " + std::to_string(comp + 1) + " compartments, ncompts is " + std::to_string(n); return false; }
    if (c != m.nchips) { err = "chip[] length does not match the ops[] walk"; return false; }
    if (P.giv && gcur > m.ngivals) { err = "givals[] shorter than the ops[] walk writes"; return false; }
    for (int64_t k = 0; k < nconcs; k++)
        if (pool_chip[k] >= 0 && pool_owner[k] < 0) refuse(P, -1, CACONC_OP, -1, HSH_R_POOL);

    // previous NEWVOLT compartment in op order (row-cache live-in)
    P.pred.assign(n, -1);
    {
        int last = -1;
        for (int i = 0; i < n; i++) {
            if (member_start(i)) last = -1;    // each solver's row cache starts empty
            if (has_volt[i]) { P.pred[i] = last; last = i; }
        }
    }

    // ------------------------------------------------------------------
    // tree analyser
    P.parent.assign(n, -1);
    P.ta.assign(n, 0.0); P.tcf.assign(n, 0.0); P.tcb.assign(n, 0.0);
    P.ra_fwd.assign(n, -1); P.ra_bwd.assign(n, -1);
    std::vector<char> got_bwd(n, 0);
    std::vector<int32_t> lastkid(n, -1);
    {
        // One solver's funcs[]/ravals[] (hines_solve.c:223-296) for rows
        // r0 .. r0+nn-1, its ravals at q0.. of the merged ravals.
        // A program with sibling ops (symmetric compartments,
        // hines_solve.c:180-184, :262-268, :289-291) runs as a unit; it is
        // checked here so the device interpreter stays inside its arrays:
        // every ravals index and row in range, the row cursor on the last row
        // at the forward FINISH, one CALC_RESULTS per row below it.  Returns
        // 2 for such a program (unit recorded), 0 if it holds no sibling op,
        // -2 if it is not one the interpreter can run.
        auto sibling_unit = [&](const int32_t* f, int64_t nf, const double* ra, int64_t nra, int nn, int r0,
                                int64_t q0) -> int {
            bool sib = false;
            int64_t fi = 0, ri = 0;
            int row = 1;
            auto idx_ok = [&](int64_t v) { return v >= 0 && v < nra; };
            auto row_ok = [&](int64_t v) { return v >= 0 && !(v & 1) && v / 2 < nn; };
            for (;;) {
                if (fi >= nf) return -2;
                const int32_t op = f[fi++];
                if (op == F_ELIM) {
                    if (fi >= nf || !row_ok(f[fi]) || ri + 2 > nra) return -2;
                    fi++;
                    ri += 2;
                } else if (op == F_SET_DIAG || op == F_SKIP_DIAG) {
                    row += op == F_SET_DIAG ? 1 : 2;
                    if (row >= nn) return -2;
                } else if (op == F_FASTSIB || op == F_COPY) {
                    if (fi >= nf || !idx_ok(f[fi]) || ri + 1 > nra) return -2;
                    fi++;
                    ri++;
                    sib = true;
                } else if (op == F_SIB) {
                    if (fi + 1 >= nf || !idx_ok(f[fi]) || !idx_ok(f[fi + 1])) return -2;
                    fi += 2;
                    sib = true;
                } else if (op == F_FINISH) {
                    break;
                } else {
                    return -2;
                }
            }
            if (row != nn - 1) return sib ? -2 : 0;
            int vi = nn - 1;
            for (;;) {
                if (fi >= nf) return -2;
                const int32_t op = f[fi++];
                if (op == F_ELIM) {
                    if (fi >= nf || !row_ok(f[fi]) || ri + 1 > nra) return -2;
                    fi++;
                    ri++;
                } else if (op == F_CALC) {
                    if (vi == 0) return -2;
                    vi--;
                } else if (op == F_SIB) {
                    if (fi + 1 >= nf || !idx_ok(f[fi]) || !row_ok(f[fi + 1])) return -2;
                    fi += 2;
                    sib = true;
                } else if (op == F_FINISH) {
                    break;
                } else {
                    return -2;
                }
            }
            if (!sib) return 0;
            if (vi != 0) return -2;
            Plan::FuncsUnit u;
            u.r0 = r0;
            u.nn = nn;
            u.f0 = (int64_t)P.ffuncs.size();
            u.nf = fi;
            u.q0 = (int64_t)P.fravals.size();
            u.nq = nra;
            u.qg = q0;
            P.ffuncs.insert(P.ffuncs.end(), f, f + fi);
            P.fravals.insert(P.fravals.end(), ra, ra + nra);
            P.funits.push_back(u);
            // the unit's rows as one tree of the layout: a chain of
            // placeholder parents ending at its last row
            for (int r = 0; r + 1 < nn; r++) P.parent[r0 + r] = r0 + r + 1;
            return 2;
        };
        auto decode = [&](const int32_t* f, int64_t nf, const double* ra, int64_t nra, int nn, int r0,
                          int64_t q0) -> int {   // 1 ok, 0 refused, -1 malformed
            int64_t fi = 0, ri = 0;
            if (nn <= 1) return 1;
            {
                const int u = sibling_unit(f, nf, ra, nra, nn, r0, q0);
                if (u == 2) return 1;
                if (u == -2) {
                    refuse(P, r0, F_SIB, -1, HSH_R_FUNCS);
                    return 0;
                }
            }
            int cur = 1;
            for (;;) {
                if (fi >= nf) { err = "funcs[] missing forward FINISH"; return -1; }
                int32_t op = f[fi++];
                if (op == F_ELIM) {
                    if (fi >= nf || ri + 2 > nra) { err = "funcs[] truncated"; return -1; }
                    int32_t kr = f[fi++];
                    if (kr < 0 || (kr & 1) || kr / 2 >= nn || kr / 2 >= cur) {
                        err = "FORWARD_ELIM row out of range"; return -1;
                    }
                    int kid = kr / 2;
                    if (P.parent[r0 + kid] != -1) { refuse(P, r0 + kid, op, (int)(fi - 2), HSH_R_TREE); return 0; }
                    if (kid <= lastkid[r0 + cur]) { refuse(P, r0 + cur, op, (int)(fi - 2), HSH_R_TREE); return 0; }
                    lastkid[r0 + cur] = kid;
                    P.parent[r0 + kid] = r0 + cur;
                    P.ra_fwd[r0 + kid] = (int32_t)(q0 + ri);
                    P.ta[r0 + kid] = ra[ri++];
                    P.tcf[r0 + kid] = ra[ri++];
                } else if (op == F_SET_DIAG) {
                    cur += 1;
                } else if (op == F_SKIP_DIAG) {
                    cur += 2;
                } else if (op == F_FINISH) {
                    break;
                } else {
                    refuse(P, r0 + cur, op, (int)(fi - 1), HSH_R_FUNCS);
                    return 0;
                }
                if (cur >= nn) { err = "funcs[] forward pass runs past the last row"; return -1; }
            }
            int row = nn - 2;
            for (;;) {
                if (fi >= nf) { err = "funcs[] missing backward FINISH"; return -1; }
                int32_t op = f[fi++];
                if (op == F_ELIM) {
                    if (fi >= nf || ri + 1 > nra || row < 0) { err = "funcs[] truncated"; return -1; }
                    int32_t pr = f[fi++];
                    if (pr != 2 * (P.parent[r0 + row] - r0) || got_bwd[r0 + row]) {
                        refuse(P, r0 + row, op, (int)(fi - 2), HSH_R_TREE);
                        return 0;
                    }
                    got_bwd[r0 + row] = 1;
                    P.ra_bwd[r0 + row] = (int32_t)(q0 + ri);
                    P.tcb[r0 + row] = ra[ri++];
                } else if (op == F_CALC) {
                    row--;
                } else if (op == F_FINISH) {
                    break;
                } else {
                    refuse(P, r0 + row, op, (int)(fi - 1), HSH_R_FUNCS);
                    return 0;
                }
            }
            if (row != -1) { err = "funcs[] backward pass does not cover every row"; return -1; }
            for (int r = 0; r < nn - 1; r++)
                if (P.parent[r0 + r] >= 0 && !got_bwd[r0 + r]) {
                    refuse(P, r0 + r, 0, -1, HSH_R_TREE);
                    return 0;
                }
            return 1;
        };
        bool ok = true;
        if (!g) {
            const int rc = decode(m.funcs, m.nfuncs, m.ravals, m.nravals, n, 0, 0);
            if (rc < 0) return false;
            ok = rc == 1;
        } else {
            for (size_t mi = 0; mi < g->row0.size() && ok; mi++) {
                const int rc = decode(g->funcs[mi], g->nfuncs[mi], g->ravals[mi], g->nravals[mi], g->nrows[mi],
                                      g->row0[mi], g->ra0[mi]);
                if (rc < 0) return false;
                ok = rc == 1;
            }
        }
        if (ok) {
            // trees are contiguous row blocks ending at their root
            std::vector<int32_t> root(n);
            for (int r = n - 1; r >= 0; r--) root[r] = (P.parent[r] < 0) ? r : root[P.parent[r]];
            int lo = 0;
            for (int r = 0; r < n; r++) {
                if (root[r] != r) {
                    if (root[r] < r) { ok = false; refuse(P, r, 0, -1, HSH_R_TREE); break; }
                    continue;
                }
                for (int q = lo; q <= r; q++)
                    if (root[q] != r) { ok = false; refuse(P, q, 0, -1, HSH_R_TREE); break; }
                if (!ok) break;
                P.tree_lo.push_back(lo);
                P.tree_hi.push_back(r);
                lo = r + 1;
            }
            std::vector<int32_t> unit_at(n, -1);
            for (size_t u = 0; u < P.funits.size(); u++) unit_at[P.funits[u].r0] = (int32_t)u;
            P.tree_funit.assign(P.tree_lo.size(), -1);
            for (size_t t = 0; t < P.tree_lo.size(); t++) P.tree_funit[t] = unit_at[P.tree_lo[t]];
        }
    }
    if (!P.refusals.empty()) return true;

    // program array: templates concatenated; program = [nslots, template id,
    // micro-ops...] (the slot count lets a wave fetch the compartment's whole
    // slot vector before interpreting it)
    P.tprog.assign(P.templates.size(), 0);
    for (size_t t = 0; t < P.templates.size(); t++) {
        P.tprog[t] = (int32_t)P.prog.size();
        P.prog.push_back(P.templates[t].nslots);
        P.prog.push_back((int32_t)t);
        P.prog.insert(P.prog.end(), P.templates[t].code.begin(), P.templates[t].code.end());
    }
    P.max_tree_rows = 0;
    for (size_t t = 0; t < P.tree_lo.size(); t++)
        P.max_tree_rows = std::max(P.max_tree_rows, P.tree_hi[t] - P.tree_lo[t] + 1);

    // algorithmic bytes of one step (fp64, single pass): Vm r+w, gate states
    // r+w, read-only slots, tree coefficients + parent for non-root rows.
    double bytes = 0;
    for (int r = 0; r < n; r++) {
        const Template& t = P.templates[P.comp_tmpl[r]];
        bytes += 16.0 + 16.0 * t.ngates + 8.0 * (t.nslots - t.ngates - t.nscratch) + 8.0 * t.nislots;
        if (P.parent[r] >= 0) bytes += 28.0;
    }
    P.bytes_per_step = bytes;
    return build_layout(P, LayoutOpts{}, err);
}

// ----------------------------------------------------------------------
// Layout.  Trees of up to o.tile_max_rows rows (and trees too large for a
// pack) go to 64-neuron tiles, lane = neuron: neurons with the same size and
// the same per-row programs share a tile, so block r of a tile is row r of 64
// neurons.  Larger trees go to packs of about o.pack_cap compartments, lane =
// compartment: the pack's compartments are grouped by program into blocks,
// and the pack's elimination runs from a level schedule.
bool build_layout(Plan& P, const LayoutOpts& o, std::string& err)
{
    const int n = (int)P.comp_tmpl.size();
    P.layout = o;
    P.blocks.clear();
    P.tiles.clear();
    P.packs.clear();
    P.lvl.clear();
    P.fent.clear();
    P.kidx.clear();
    P.kta.clear();
    P.ktcf.clear();
    P.bkp.clear();
    P.btcb.clear();
    P.max_rows = 0;
    P.max_pack_pos = 0;
    P.pack_lds = 0;
    P.pack_lds_stage = 0;
    P.pack_lds_stage_idx = 0;
    P.comp_pos.assign(n, -1);
    const int ntrees = (int)P.tree_lo.size();
    auto rows_of = [&](int t) { return P.tree_hi[t] - P.tree_lo[t] + 1; };
    // a pack's level schedule keeps positions in 15 bits, and its V/rhs/diag
    // and predecessor index (28 bytes a position) must fit the LDS
    constexpr int PACK_MAX_POS = 4096;
    const auto is_unit = [&](int t) { return !P.tree_funit.empty() && P.tree_funit[t] >= 0; };
    auto packable = [&](int t) {
        const int R = rows_of(t);
        if (R <= o.tile_max_rows || is_unit(t)) return false;
        // worst case: every row of the tree alone in its program's block
        const int ntmpl = (int)P.templates.size();
        return (int64_t)R + (int64_t)ntmpl * (W - 1) <= PACK_MAX_POS;
    };

    std::map<std::vector<int32_t>, std::vector<int32_t>> groups;   // signature -> trees
    std::vector<int32_t> big;
    for (int t = 0; t < ntrees; t++) {
        if (packable(t)) {
            big.push_back(t);
            continue;
        }
        std::vector<int32_t> sig;
        sig.push_back(rows_of(t));
        for (int r = P.tree_lo[t]; r <= P.tree_hi[t]; r++) sig.push_back(P.comp_tmpl[r]);
        sig.push_back(is_unit(t) ? 1 : 0);     // units share tiles only with units
        groups[sig].push_back(t);
    }
    // larger trees first: they run longest in the persistent kernel; the
    // units' tiles after all others (the tile kernels take the first
    // ntiles_regular tiles)
    std::vector<const std::pair<const std::vector<int32_t>, std::vector<int32_t>>*> order;
    for (auto& g : groups) order.push_back(&g);
    std::stable_sort(order.begin(), order.end(), [](auto* a, auto* b) {
        if (a->first.back() != b->first.back()) return a->first.back() < b->first.back();
        return a->first[0] > b->first[0];
    });
    P.funit_tile.assign(P.funits.size(), -1);
    P.funit_lane.assign(P.funits.size(), -1);
    P.ntiles_regular = 0;

    int64_t nS = 0, nI = 0;
    auto add_block = [&](int ti, int tile, int lanes) {
        BlockMeta bm{};
        bm.prog = P.tprog[ti];
        bm.dslot = (int32_t)nS;
        bm.islot = (int32_t)nI;
        bm.tile = tile;
        bm.tmpl = ti;
        bm.lanes = lanes;
        P.blocks.push_back(bm);
        nS += (int64_t)P.templates[ti].nslots * W;
        nI += (int64_t)P.templates[ti].nislots * W;
        return (int)P.blocks.size() - 1;
    };
    for (auto* g : order) {
        const std::vector<int32_t>& sig = g->first;
        const std::vector<int32_t>& trees = g->second;
        int R = sig[0];
        const bool units = sig.back() != 0;
        if (!units) P.max_rows = std::max(P.max_rows, R);
        for (size_t t0 = 0; t0 < trees.size(); t0 += W) {
            TileMeta tm;
            tm.block0 = (int32_t)P.blocks.size();
            tm.rows = R;
            tm.lanes = (int32_t)std::min<size_t>(W, trees.size() - t0);
            tm.units = units ? 1 : 0;
            int tile = (int)P.tiles.size();
            P.tiles.push_back(tm);
            if (!units) P.ntiles_regular = (int)P.tiles.size();
            for (int l = 0; units && l < tm.lanes; l++) {
                P.funit_tile[P.tree_funit[trees[t0 + l]]] = tile;
                P.funit_lane[P.tree_funit[trees[t0 + l]]] = l;
            }
            for (int r = 0; r < R; r++) {
                const int b = add_block(sig[1 + r], tile, tm.lanes);
                for (int l = 0; l < tm.lanes; l++) {
                    int row = P.tree_lo[trees[t0 + l]] + r;
                    P.comp_pos[row] = b * W + l;
                }
            }
        }
    }

    // packs: largest trees first, consecutive trees of one size together
    std::stable_sort(big.begin(), big.end(), [&](int a, int b) { return rows_of(a) > rows_of(b); });
    // height (leaves 0), depth (roots 0) and children (ascending row order:
    // the reference eliminates kids in that order) of every row
    std::vector<int32_t> height(n, 0), depth(n, 0), kid_off(n + 1, 0), kid_list;
    for (int r = 0; r < n; r++)
        if (P.parent[r] >= 0) {
            height[P.parent[r]] = std::max(height[P.parent[r]], height[r] + 1);   // kids precede parents
            kid_off[P.parent[r] + 1]++;
        }
    for (int r = n - 1; r >= 0; r--) depth[r] = P.parent[r] >= 0 ? depth[P.parent[r]] + 1 : 0;
    for (int r = 0; r < n; r++) kid_off[r + 1] += kid_off[r];
    kid_list.assign(kid_off[n], 0);
    {
        std::vector<int32_t> cur(kid_off.begin(), kid_off.end() - 1);
        for (int r = 0; r < n; r++)
            if (P.parent[r] >= 0) kid_list[cur[P.parent[r]]++] = r;
    }
    const int ntiles = (int)P.tiles.size();
    const int cap = std::max(1, o.pack_cap);
    for (size_t i0 = 0; i0 < big.size();) {
        // trees of this pack
        size_t i1 = i0;
        int tot = 0;
        const int ntmpl = (int)P.templates.size();
        while (i1 < big.size()) {
            const int R = rows_of(big[i1]);
            if (i1 > i0 && (tot + R > cap || (int64_t)tot + R + (int64_t)ntmpl * (W - 1) > PACK_MAX_POS)) break;
            tot += R;
            i1++;
        }
        PackMeta pm{};
        pm.block0 = (int32_t)P.blocks.size();
        const int pk = (int)P.packs.size();
        // blocks by program: compartments in tree order, rows ascending
        std::vector<std::vector<int32_t>> by_t(ntmpl);
        for (size_t i = i0; i < i1; i++)
            for (int r = P.tree_lo[big[i]]; r <= P.tree_hi[big[i]]; r++) by_t[P.comp_tmpl[r]].push_back(r);
        for (int ti = 0; ti < ntmpl; ti++) {
            const auto& rows = by_t[ti];
            for (size_t c0 = 0; c0 < rows.size(); c0 += W) {
                const int lanes = (int)std::min<size_t>(W, rows.size() - c0);
                const int b = add_block(ti, ntiles + pk, lanes);
                for (int l = 0; l < lanes; l++) P.comp_pos[rows[c0 + l]] = b * W + l;
            }
        }
        pm.nblk = (int32_t)P.blocks.size() - pm.block0;
        const int base = pm.block0 * W;
        const int NP = pm.nblk * W;
        if (NP > PACK_MAX_POS) { err = "internal: pack exceeds its position limit"; return false; }
        P.max_pack_pos = std::max(P.max_pack_pos, NP);
        // level schedule
        int H = 0, D = 0;
        for (size_t i = i0; i < i1; i++)
            for (int r = P.tree_lo[big[i]]; r <= P.tree_hi[big[i]]; r++) {
                H = std::max(H, height[r]);
                D = std::max(D, depth[r]);
            }
        auto lp = [&](int32_t r) { return P.comp_pos[r] - base; };
        std::vector<std::vector<int32_t>> flev(H + 1), blev(D + 1);
        for (size_t i = i0; i < i1; i++)
            for (int r = P.tree_lo[big[i]]; r <= P.tree_hi[big[i]]; r++) {
                if (height[r] > 0) flev[height[r]].push_back(r);
                blev[depth[r]].push_back(r);
            }
        auto by_lp = [&](int32_t a, int32_t b) { return lp(a) < lp(b); };
        pm.lv0 = (int32_t)P.lvl.size();
        pm.nfl = H;
        pm.nbl = D + 1;
        pm.fe0 = (int32_t)(P.fent.size() / 2);
        pm.ke0 = (int32_t)P.kidx.size();
        pm.be0 = (int32_t)P.bkp.size();
        int fe = 0, ke = 0, be = 0;
        for (int h = 1; h <= H; h++) {
            P.lvl.push_back(fe);
            std::sort(flev[h].begin(), flev[h].end(), by_lp);
            for (int32_t pr : flev[h]) {
                const int32_t nk = kid_off[pr + 1] - kid_off[pr];
                P.fent.push_back(lp(pr) | (nk << 16));
                P.fent.push_back(ke);
                for (int32_t q = kid_off[pr]; q < kid_off[pr + 1]; q++) {
                    const int32_t c = kid_list[q];
                    P.kidx.push_back((uint16_t)lp(c));
                    P.kta.push_back(P.ta[c]);
                    P.ktcf.push_back(P.tcf[c]);
                    ke++;
                }
                fe++;
            }
        }
        P.lvl.push_back(fe);
        for (int dd = 0; dd <= D; dd++) {
            P.lvl.push_back(be);
            std::sort(blev[dd].begin(), blev[dd].end(), by_lp);
            for (int32_t r : blev[dd]) {
                const int32_t pr = P.parent[r];
                P.bkp.push_back(lp(r) | ((pr >= 0 ? lp(pr) + 1 : 0) << 16));
                P.btcb.push_back(pr >= 0 ? P.tcb[r] : 0.0);
                be++;
            }
        }
        P.lvl.push_back(be);
        pm.nfe = fe;
        pm.nke = ke;
        pm.nbe = be;
        pm.last_lp = -1;
        P.packs.push_back(pm);
        // LDS of the pack kernel: V, rhs, diag (+ predecessor index), and the
        // staged schedule
        const int64_t lds = (int64_t)NP * (3 * 8 + 4);
        const int64_t stage = (int64_t)ke * (8 + 8 + 2) + (int64_t)be * (8 + 4) + (int64_t)fe * 8 +
                              (int64_t)(H + D + 3) * 4 + 16;
        P.pack_lds = std::max<int64_t>(P.pack_lds, lds);
        P.pack_lds_stage = std::max<int64_t>(P.pack_lds_stage, lds + stage);
        const int64_t stage_idx = (int64_t)ke * 2 + (int64_t)be * 4 + (int64_t)fe * 8 + (int64_t)(H + D + 3) * 4 + 16;
        P.pack_lds_stage_idx = std::max<int64_t>(P.pack_lds_stage_idx, lds + stage_idx);
        i0 = i1;
    }
    for (auto& pm : P.packs) {
        const int last = P.comp_pos[n - 1];
        if (last >= pm.block0 * W && last < (pm.block0 + pm.nblk) * W) pm.last_lp = last - pm.block0 * W;
    }
    if (nS > INT32_MAX || nI > INT32_MAX || (int64_t)P.blocks.size() * W > INT32_MAX) {
        err = "model too large for 32-bit slot offsets";
        return false;
    }
    P.nS = nS;
    P.nI = nI;
    P.pos_comp.assign(P.blocks.size() * W, -1);
    for (int r = 0; r < n; r++) P.pos_comp[P.comp_pos[r]] = r;
    // padding lanes duplicate lane 0 of their block
    for (size_t b = 0; b < P.blocks.size(); b++)
        for (int l = 1; l < W; l++)
            if (P.pos_comp[b * W + l] < 0) P.pos_comp[b * W + l] = P.pos_comp[b * W];
    return true;
}

}  // namespace hsh
