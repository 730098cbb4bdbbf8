// hsh_jit.cpp -- run-time specialisation of the channel programs.
//
// The interpreter (hsh_device.h run_prog) pays, per micro-op, a scalar load,
// address arithmetic and a branch tree, and it cannot overlap the table
// gathers of different gates.  A model has few distinct programs (the
// 50,000 x 16 workload has two), so at hsh_create the library writes one
// straight-line device function per program -- constant slot offsets, table
// columns and powers, the same helper functions and expression order as the
// interpreter -- compiles it with hipRTC for the device's architecture and
// launches the same kernel bodies with it.  Compilation failures fall back to
// the built-in interpreter kernels (still on the device).
#include <hip/hiprtc.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "hsh_internal.h"

namespace hsh {

namespace {

constexpr int MAXG_ADJ = 3;   // gates per channel the adjoint runs (hsh_adjoint_device.h MAXG)

#include "hsh_jit_src.inc"   // kDevTypesSrc, kDeviceSrc, kAdjointSrc: hsh_devtypes.h, hsh_device.h,
                             // hsh_adjoint_device.h

std::mutex g_jit_mu;
std::map<std::string, std::vector<char>> g_code_cache;   // source -> code object

std::string slot(int c) { return "s" + std::to_string(c); }

std::string gen_template(const Template& t, int id, int ncols, bool giv)
{
    std::string o;
    o += "__device__ __forceinline__ CompOut tmpl_" + std::to_string(id) +
         "(const DevModel& d, double* __restrict__ Sb, int* __restrict__ Ib, int lane, double Vm, "
         "bool pred_ok, double Vp, bool active, int step)\n{\n";
    for (int c = 0; c < t.nslots; c++)
        o += "    const double " + slot(c) + " = Sb[" + std::to_string(c) + " * W + lane];\n";
    o += "    double sumgchan = 0.0;\n    double ichan = s0 + s1;\n";
    o += "    double Gk = 0.0, Ek = 0.0, Ik = 0.0, vipol = 0.0, cipol = 0.0;\n";
    o += "    double flux[MAXPOOL] = {0.0, 0.0, 0.0, 0.0};\n    double csum = 0.0; (void)csum;\n";
    // The diffusion systems' per-step arrays (codiag and coresult per shell,
    // elimination scratch) live in the thread's own arrays here: codiag is 1
    // and coresult 0 between steps, so they start from those values.
    std::map<int, std::string> tmap;      // slot -> its element of those arrays
    {
        int k = 0;
        for (size_t q = 0; q < t.code.size();) {
            const int op = t.code[q++];
            int skip = 0;
            switch (op) {
            case U_END: q = t.code.size(); break;
            case U_IPOL1V: case U_SYN2: case U_SYN2MG: case U_GIV0: case U_TABGHK: case U_TABCURR: case U_ELMMPUMP:
            case U_FLUXCONC: case U_CTAUPUMP: case U_TAUPUMP: case U_NERNST: case U_GHK: skip = 2; break;
            case U_SPIKE: case U_CONCVAL: case U_NEWCONC1: case U_NEWCONC2: case U_FLUX: case U_SCALEFLUX:
            case U_CSTORE: case U_POOL: skip = 1; break;
            case U_IPOLG: case U_POOLCA: case U_MMPUMP: case U_BUFFER: skip = 3; break;
            case U_HILLPUMP: case U_PHILLPUMP: skip = 4; break;
            case U_DIFF: {
                const int ms = t.code[q], dw = t.code[q + 1], N = ms * dw, nx = dw == 1 ? ms : N * dw;
                const int dg = t.code[q + 6], cr = t.code[q + 7], x = t.code[q + 8], r = t.code[q + 9];
                const std::string K = std::to_string(k);
                o += "    double T" + K + "d[" + std::to_string(ms) + "], T" + K + "c[" + std::to_string(ms) +
                     "], T" + K + "x[" + std::to_string(nx) + "], T" + K + "r[" + std::to_string(N) + "];\n";
                o += "    for (int i = 0; i < " + std::to_string(ms) + "; i++) { T" + K + "d[i] = 1.0; T" + K +
                     "c[i] = 0.0; }\n";
                for (int i = 0; i < ms; i++) {
                    tmap[dg + i] = "T" + K + "d[" + std::to_string(i) + "]";
                    tmap[cr + i] = "T" + K + "c[" + std::to_string(i) + "]";
                }
                for (int i = 0; i < nx; i++) tmap[x + i] = "T" + K + "x[" + std::to_string(i) + "]";
                for (int i = 0; i < N; i++) tmap[r + i] = "T" + K + "r[" + std::to_string(i) + "]";
                k++;
                skip = 10;
                break;
            }
            default: break;
            }
            q += skip;
        }
    }
    auto tref = [&](int slot) {
        auto it = tmap.find(slot);
        return it != tmap.end() ? it->second : "Sb[" + std::to_string(slot) + " * W + lane]";
    };
    o += "    const double* vtab = d.tab;\n    const double* ctab = d.tab;\n    const double* tabr = d.tab;\n";
    o += "    int cfilo = 0;\n    double ipol = 0.0, ghk = 0.0;\n";
    o += "    (void)Ib; (void)active; (void)step; (void)Gk; (void)Ek; (void)Ik; (void)vipol; (void)cipol;\n";
    o += "    (void)flux; (void)vtab; (void)ctab; (void)tabr; (void)cfilo; (void)ipol; (void)ghk;\n";
    o += "    (void)pred_ok; (void)Vp;\n";
    // givals (only in modules made for a model that asks for them)
    const std::string GM = giv ? "gmax = " : "";
    const std::string PUT3 = giv ? " if (giv) { giv_put3(d, gi, Gk, Ik, Ek, gmax); Im += Ik; }" : "";
    if (giv) o += "    const bool giv = active;\n    int gi = 0;\n    double Im = 0.0, gmax = 0.0;\n"
                  "    (void)gmax; (void)Im;\n";
    const std::string NC = std::to_string(ncols);
    int c = 2;
    size_t pc = 0;
    const std::vector<int32_t>& code = t.code;
    while (pc < code.size()) {
        const int op = code[pc++];
        switch (op) {
        case U_NEWVOLT:
            o += "    { int filo; newvolt(d, Vm, pred_ok, Vp, filo, vipol); vtab = d.tab + (long long)filo * " +
                 NC + "; }\n";
            break;
        case U_CHAN_EK:
            o += "    Ek = " + slot(c) + "; " + GM + "Gk = " + slot(c + 1) + ";\n";
            c += 2;
            break;
        case U_CHAN:
            o += "    " + GM + "Gk = " + slot(c) + ";\n";
            c += 1;
            break;
        case U_GIV0: {
            const int rs = code[pc], is = code[pc + 1];
            pc += 2;
            if (giv)
                o += "    if (giv) { gi = Ib[" + std::to_string(is) + " * W + lane]; Im = s0 - Vm / " + slot(rs) +
                     "; d.giv[gi + 1] = Im; gi += 2; }\n";
            break;
        }
        case U_IPOL1V: {
            const int col = code[pc], k = code[pc + 1];
            pc += 2;
            o += "    { const double X = gate_update(vtab, " + std::to_string(col) + ", " + NC +
                 ", vipol, " + slot(c) + ", " + (k < 0 ? "true" : "false") + "); tabr = vtab + " +
                 std::to_string(col + 1) + "; Sb[" + std::to_string(c) +
                 " * W + lane] = X; Gk = apply_power(Gk, X, " + std::to_string(k < 0 ? -k : k) + "); }\n";
            c += 1;
            break;
        }
        case U_ADDCURR:
            o += "    sumgchan += Gk; ichan += Ek * Gk; Ik = Gk * (Ek - Vm);" + PUT3 + "\n";
            break;
        case U_SPIKE: {
            const int is = code[pc++];
            o += "    spike_op(d, Ib, " + std::to_string(is) + ", lane, Vm, " + slot(c) +
                 ", active, step);\n";
            c += 1;
            break;
        }
        case U_SYN2:
        case U_SYN2MG: {
            const int k = code[pc], is = code[pc + 1];
            pc += 2;
            o += "    Gk *= syn2_op(d, Sb, Ib, " + std::to_string(c) + ", " + std::to_string(k) + ", " +
                 std::to_string(is) + ", lane, " + slot(c) + ", " + slot(c + 1) + ", active, step);";
            c += 2;
            if (op == U_SYN2MG) {
                if (giv) o += " if (giv) giv_put3(d, gi, Gk, Gk * (Ek - Vm), Ek, gmax);";
                o += " Gk = mgblock(Gk, Vm, " + slot(c) + ", " + slot(c + 1) + ", " + slot(c + 2) + ");";
                c += 3;
            }
            o += " sumgchan += Gk; ichan += Ek * Gk; Ik = Gk * (Ek - Vm);" + PUT3 + "\n";
            break;
        }
        case U_CONCVAL:
            o += "    flux[" + std::to_string(code[pc++]) + "] = 0.0;\n";
            break;
        case U_NEWCONC1:
        case U_DOMAIN: {
            std::string conc;
            if (op == U_NEWCONC1) {
                conc = slot(code[pc++]);
            } else {
                conc = slot(c) + " * Ik";
                c += 1;
            }
            o += "    newconc(d, " + conc + ", cfilo, cipol); ctab = d.tab + (long long)cfilo * " + NC + ";\n";
            break;
        }
        case U_NEWCONC2:
            o += "    tabr = ctab; ipol = cipol; newconc(d, " + slot(code[pc++]) +
                 ", cfilo, cipol); ctab = d.tab + (long long)cfilo * " + NC + ";\n";
            break;
        case U_IPOLG: {
            const int which = code[pc], col = code[pc + 1], k = code[pc + 2];
            pc += 3;
            o += "    { double X, Y; ipol_xy(d, " + std::to_string(which) + ", " + std::to_string(col) +
                 ", vtab, ctab, tabr, cfilo, vipol, cipol, ipol, X, Y); X = " +
                 (k > 0 ? "(" + slot(c) + " * (2.0 - X) + Y) / X" : std::string("Y / X")) + "; Sb[" +
                 std::to_string(c) + " * W + lane] = X; Gk = apply_power(Gk, X, " +
                 std::to_string(k < 0 ? -k : k) + "); }\n";
            c += 1;
            break;
        }
        case U_TABGHK: {
            const int col = code[pc], which = code[pc + 1];
            pc += 2;
            o += "    { double X, Y; ipol_xy(d, " + std::to_string(which) + ", " + std::to_string(col) +
                 ", vtab, ctab, tabr, cfilo, vipol, cipol, ipol, X, Y); Ek = Y; ghk = X; }\n";
            break;
        }
        case U_SAMEGHK:
            o += "    " + (giv ? "gmax = " + slot(c) + "; " : std::string()) + "Gk = " + slot(c) + " * ghk;\n";
            c += 1;
            break;
        case U_IPOL1C: {
            const int col = code[pc], k = code[pc + 1];
            pc += 2;
            o += "    { const double X = gate_update(ctab, " + std::to_string(col) + ", " + NC +
                 ", cipol, " + slot(c) + ", " + (k < 0 ? "true" : "false") + "); tabr = ctab + " +
                 std::to_string(col + 1) + "; Sb[" + std::to_string(c) +
                 " * W + lane] = X; Gk = apply_power(Gk, X, " + std::to_string(k < 0 ? -k : k) + "); }\n";
            c += 1;
            break;
        }
        case U_TABCURR: {
            const int col = code[pc], which = code[pc + 1];
            pc += 2;
            o += "    { double X, Y; ipol_xy(d, " + std::to_string(which) + ", " + std::to_string(col) +
                 ", vtab, ctab, tabr, cfilo, vipol, cipol, ipol, X, Y); Ek = Y; Gk *= X; }"
                 " sumgchan += Gk; ichan += Ek * Gk; Ik = Gk * (Ek - Vm);" + PUT3 + "\n";
            break;
        }
        case U_FLUX:
            o += "    flux[" + std::to_string(code[pc++]) + "] += Ik;\n";
            break;
        case U_SCALEFLUX:
            o += "    flux[" + std::to_string(code[pc++]) + "] += " + slot(c) + " * Ik;\n";
            c += 1;
            break;
        case U_NERNST: {
            const int ci = code[pc], co = code[pc + 1];
            pc += 2;
            o += "    Ek = log(" + slot(co) + " / " + slot(ci) + ") * " + slot(c) + "; " + GM + "Gk = " +
                 slot(c + 1) + ";\n";
            c += 2;
            break;
        }
        case U_GHK: {
            const int ci = code[pc], co = code[pc + 1];
            pc += 2;
            o += "    { const GhkOut g = ghk_call(Vm, " + slot(ci) + ", " + slot(co) + ", " + slot(c) + ", " +
                 slot(c + 1) + "); Ek = g.Ek; ghk = g.slope; " +
                 (giv ? "gmax = " + slot(c + 2) + "; " : std::string()) + "Gk = " + slot(c + 2) + " * g.slope; }\n";
            c += 3;
            break;
        }
        case U_POOLCA: {
            const int p = code[pc], cs = code[pc + 1], ks = code[pc + 2];
            pc += 3;
            o += "    Sb[" + std::to_string(cs) + " * W + lane] = caconc(" + slot(cs) + ", " + slot(ks) + ", " +
                 slot(ks + 1) + ", flux[" + std::to_string(p) + "], " + slot(ks + 2) + ");\n";
            break;
        }
        case U_ELMMPUMP: {
            const int cs = code[pc], dg = code[pc + 1];
            pc += 2;
            const std::string C = "Sb[" + std::to_string(cs) + " * W + lane]";
            o += "    { const double X = 1.0 / (" + C + " + " + slot(c) + "); Ik = " + slot(c + 1) + " * " + C +
                 " * X; ichan += Ik;";
            if (giv) o += " if (giv) { d.giv[gi] = d.chanmode5 ? Ik / " + slot(c + 1) + " : Ik; gi += 1; } Im += Ik;";
            if (dg >= 0) o += " " + tref(dg) + " += " + slot(c + 2) + " * X;";
            o += " }\n";
            c += 3;
            break;
        }
        case U_CZERO:
            o += "    csum = 0.0;\n";
            break;
        case U_CSTORE:
            o += "    " + tref(code[pc++]) + " = csum;\n";
            break;
        case U_FLUXCONC: {
            const int a = code[pc], p = code[pc + 1];
            pc += 2;
            o += "    conc_fluxconc(Sb, lane, csum, " + std::to_string(a) + ", " +
                 (p >= 0 ? "flux[" + std::to_string(p) + "]" : std::string("0.0")) + ");\n";
            break;
        }
        case U_CTAUPUMP:
            o += "    conc_ctaupump(Sb, lane, csum, " + std::to_string(code[pc]) + ", " + tref(code[pc + 1]) + ");\n";
            pc += 2;
            break;
        case U_TAUPUMP:
            o += "    conc_taupump(Sb, lane, csum, " + std::to_string(code[pc]) + ", " + tref(code[pc + 1]) +
                 ", d.dt, Vm);\n";
            pc += 2;
            break;
        case U_MMPUMP:
            o += "    conc_mmpump(Sb, lane, " + std::to_string(code[pc]) + ", " + tref(code[pc + 1]) + ", " +
                 tref(code[pc + 2]) + ");\n";
            pc += 3;
            break;
        case U_HILLPUMP:
        case U_PHILLPUMP: {
            const std::string args = std::to_string(code[pc]) + ", " + tref(code[pc + 1]) + ", " +
                                     tref(code[pc + 2]) + ", " + std::to_string(code[pc + 3]);
            pc += 4;
            o += op == U_HILLPUMP ? "    conc_hillpump(Sb, lane, " + args + ");\n"
                                  : "    conc_phillpump(Sb, lane, csum, " + args + ");\n";
            break;
        }
        case U_BUFFER:
            o += "    conc_buffer(Sb, lane, csum, " + std::to_string(code[pc]) + ", " + std::to_string(code[pc + 1]) +
                 ", " + std::to_string(code[pc + 2]) + ");\n";
            pc += 3;
            break;
        case U_POOL:
            o += "    Sb[" + std::to_string(code[pc++]) + " * W + lane] += csum;\n";
            break;
        case U_DIFF: {
            const std::string dg = tref(code[pc + 6]), cr = tref(code[pc + 7]), x = tref(code[pc + 8]),
                              r = tref(code[pc + 9]);
            auto arr = [](const std::string& e) { return e.substr(0, e.find('[')); };
            o += "    diff_solve_t<1>(Sb, lane";
            for (int i = 0; i < 6; i++) o += ", " + std::to_string(code[pc + i]);
            o += ", " + arr(dg) + ", " + arr(cr) + ", " + arr(x) + ", " + arr(r) + ", d.crank);\n";
            pc += 10;
            break;
        }
        default:   // U_END
            if (giv) o += "    if (giv) d.giv[gi] = Im;\n";
            o += "    { CompOut o; o.rhs = Vm + ichan * " + slot(c) + "; o.diag = sumgchan * " + slot(c) +
                 " + " + slot(c + 1) + "; return o; }\n";
            c += 2;
            break;
        }
    }
    o += "}\n";
    return o;
}

// The reverse of one program for the adjoint's pack kernel: adj_reverse_local
// (hsh_adjoint_device.h) unrolled for this program -- the same expressions in
// the same order, so the gradients are the interpreter's bit for bit -- with
// every slot and slot adjoint loaded at the start and the adjoints stored at
// the end, so the loads of a program issue together.  Programs the adjoint
// refuses (ops past the voltage-gated set) get a stub that is never run.
std::string gen_adj_template(const Template& t, int id, int ncols)
{
    std::string o;
    o += "__device__ __forceinline__ double adj_tmpl_" + std::to_string(id) +
         "(const DevModel& d, const double* __restrict__ Sb, double* __restrict__ LSb, int lane, double Vm, "
         "double bb, double db)\n{\n";
    bool ok = true;
    int ng_chan = 0;
    for (size_t pc = 0; pc < t.code.size();) {
        const int op = t.code[pc++];
        if (op == U_IPOL1V) {
            pc += 2;
            if (++ng_chan > MAXG_ADJ) ok = false;
        } else if (op == U_CHAN || op == U_CHAN_EK) {
            ng_chan = 0;
        } else if (op == U_SPIKE) {
            pc += 1;
        } else if (op == U_NEWVOLT || op == U_ADDCURR) {
        } else if (op == U_END) {
            break;
        } else {
            ok = false;
            break;
        }
    }
    if (!ok) {
        o += "    (void)d; (void)Sb; (void)LSb; (void)lane; (void)Vm; (void)bb; (void)db;\n    return 0.0;\n}\n";
        return o;
    }
    const int ns = t.nslots;
    auto S = [](int c) { return "s" + std::to_string(c); };
    auto L = [](int c) { return "l" + std::to_string(c); };
    for (int c = 0; c < ns; c++) o += "    const double " + S(c) + " = Sb[" + std::to_string(c) + " * W + lane];\n";
    for (int c = 0; c < ns; c++) o += "    double " + L(c) + " = LSb[" + std::to_string(c) + " * W + lane];\n";
    const std::string NC = std::to_string(ncols);
    o += "    const double tbyc = " + S(ns - 2) + ";\n";
    o += "    const double lam_i = bb * tbyc;\n    const double lam_g = db * tbyc;\n";
    o += "    double vbar = bb;\n    double sumg = 0.0;\n    double ichan = s0 + s1;\n";
    o += "    l0 += lam_i;\n    l1 += lam_i;\n";
    o += "    double Gk = 0.0, Ek = 0.0, gmax = 0.0;\n    Interp ip{0, 0.0, false};\n    const double* vtab = d.tab;\n";
    o += "    (void)Gk; (void)Ek; (void)gmax; (void)ip; (void)vtab;\n";
    int c = 2, ekslot = -1, gslot = -1, gid = 0;
    std::vector<int> gs, gp, gn;      // gates of the current channel: slot, power, variable id
    for (size_t pc = 0; pc < t.code.size();) {
        const int op = t.code[pc++];
        if (op == U_IPOL1V) {
            const int col = t.code[pc], k = t.code[pc + 1];
            pc += 2;
            const std::string g = std::to_string(gid), C = std::to_string(col);
            o += "    const double* tab" + g + " = vtab + " + C + ";\n";
            o += "    const double B" + g + " = tab" + g + "[0] + ip.vipol * (tab" + g + "[" + NC + "] - tab" + g + "[0]);\n";
            o += "    const double dB" + g + " = ip.inside ? (tab" + g + "[" + NC + "] - tab" + g + "[0]) * d.invdx : 0.0;\n";
            o += "    const double dA" + g + " = ip.inside ? (tab" + g + "[" + NC + " + 1] - tab" + g +
                 "[1]) * d.invdx : 0.0;\n";
            if (k < 0) {
                o += "    const double X" + g + " = gate_update(vtab, " + C + ", " + NC + ", ip.vipol, " + S(c) + ", true);\n";
                o += "    const double gdx" + g + " = 0.0;\n";
                o += "    const double gdv" + g + " = (dA" + g + " - X" + g + " * dB" + g + ") / B" + g + ";\n";
            } else {
                o += "    const double X" + g + " = gate_update(vtab, " + C + ", " + NC + ", ip.vipol, " + S(c) + ", false);\n";
                o += "    const double gdx" + g + " = (2.0 - B" + g + ") / B" + g + ";\n";
                o += "    const double gdv" + g + " = (dA" + g + " - " + S(c) + " * dB" + g + " - X" + g + " * dB" + g +
                     ") / B" + g + ";\n";
            }
            const int pw = k < 0 ? -k : k;
            o += "    Gk = apply_power(Gk, X" + g + ", " + std::to_string(pw) + ");\n";
            gs.push_back(c);
            gp.push_back(pw);
            gn.push_back(gid);
            gid++;
            c++;
        } else if (op == U_CHAN_EK) {
            ekslot = c;
            gslot = c + 1;
            o += "    Ek = " + S(c) + ";\n    gmax = Gk = " + S(c + 1) + ";\n";
            c += 2;
            gs.clear(); gp.clear(); gn.clear();
        } else if (op == U_CHAN) {
            gslot = c;
            o += "    gmax = Gk = " + S(c) + ";\n";
            c += 1;
            gs.clear(); gp.clear(); gn.clear();
        } else if (op == U_ADDCURR) {
            o += "    {\n        sumg += Gk;\n        ichan += Ek * Gk;\n        const double lamG = lam_i * Ek + lam_g;\n";
            if (ekslot >= 0) o += "        " + L(ekslot) + " += lam_i * Gk;\n";
            o += "        double prod = 1.0;\n";
            for (size_t q = 0; q < gs.size(); q++)
                o += "        prod *= ipow(X" + std::to_string(gn[q]) + ", " + std::to_string(gp[q]) + ");\n";
            if (gslot >= 0) o += "        " + L(gslot) + " += lamG * prod;\n";
            for (size_t q = 0; q < gs.size(); q++) {
                o += "        {\n            double dG = gmax * (double)" + std::to_string(gp[q]) + " * ipow(X" +
                     std::to_string(gn[q]) + ", " + std::to_string(gp[q] - 1) + ");\n";
                for (size_t h = 0; h < gs.size(); h++)
                    if (h != q)
                        o += "            dG *= ipow(X" + std::to_string(gn[h]) + ", " + std::to_string(gp[h]) + ");\n";
                o += "            const double lamX = " + L(gs[q]) + " + lamG * dG;\n";
                o += "            " + L(gs[q]) + " = lamX * gdx" + std::to_string(gn[q]) + ";\n";
                o += "            vbar += lamX * gdv" + std::to_string(gn[q]) + ";\n        }\n";
            }
            o += "    }\n";
            gs.clear(); gp.clear(); gn.clear();
        } else if (op == U_NEWVOLT) {
            o += "    ip = table_row(d, Vm);\n    vtab = d.tab + (long long)ip.filo * " + NC + ";\n";
        } else if (op == U_SPIKE) {
            pc += 1;
            c += 1;
        } else {   // U_END
            o += "    " + L(c) + " += bb * ichan + db * sumg;\n";
            o += "    " + L(c + 1) + " += db;\n";
            break;
        }
    }
    for (int q = 0; q < ns; q++) o += "    LSb[" + std::to_string(q) + " * W + lane] = " + L(q) + ";\n";
    o += "    return vbar;\n}\n";
    return o;
}

std::string strip_pragma_once(const char* s)
{
    std::string r(s);
    const std::string p = "#pragma once";
    size_t i;
    while ((i = r.find(p)) != std::string::npos) r.erase(i, p.size());
    return r;
}

}  // namespace

std::string jit_source(const Plan& P, int ncols)
{
    std::string src;
    src += strip_pragma_once(kDevTypesSrc);
    src += strip_pragma_once(kDeviceSrc);
    src += strip_pragma_once(kAdjointSrc);
    src += "namespace hsh {\n";
    for (size_t t = 0; t < P.templates.size(); t++) src += gen_template(P.templates[t], (int)t, ncols, P.giv);
    for (size_t t = 0; t < P.templates.size(); t++) src += gen_adj_template(P.templates[t], (int)t, ncols);
    src += "struct JitAdj {\n"
           "    static __device__ __forceinline__ double rev(const DevModel& d, const BlockMeta& bm, const double* Sb, "
           "double* LSb, int lane, double Vm, double bb, double db)\n    {\n"
           "        switch (bm.tmpl) {\n";
    for (size_t t = 1; t < P.templates.size(); t++)
        src += "        case " + std::to_string(t) + ": return adj_tmpl_" + std::to_string(t) +
               "(d, Sb, LSb, lane, Vm, bb, db);\n";
    src += "        default: return adj_tmpl_0(d, Sb, LSb, lane, Vm, bb, db);\n        }\n    }\n};\n";
    src += "struct JitRun {\n"
           "    static __device__ __forceinline__ CompOut run(const DevModel& d, const BlockMeta& bm, "
           "int lane, double Vm, bool pok, double Vp, bool active, int step)\n    {\n"
           "        double* Sb = d.S + bm.dslot;\n        int* Ib = d.I + bm.islot;\n"
           "        switch (bm.tmpl) {\n";
    for (size_t t = 1; t < P.templates.size(); t++)
        src += "        case " + std::to_string(t) + ": return tmpl_" + std::to_string(t) +
               "(d, Sb, Ib, lane, Vm, pok, Vp, active, step);\n";
    src += "        default: return tmpl_0(d, Sb, Ib, lane, Vm, pok, Vp, active, step);\n"
           "        }\n    }\n};\n}  // namespace hsh\n";
    src += "extern \"C\" __global__ __launch_bounds__(256) void hsh_jit_channel(hsh::DevModel d)\n"
           "{ hsh::body_channel<hsh::JitRun>(d); }\n";
    for (int r : {1, 2, 4})
        src += "extern \"C\" __global__ __launch_bounds__(256) void hsh_jit_batched_r" + std::to_string(r) +
               "(hsh::DevModel d, int K)\n{ extern __shared__ double lds[]; hsh::body_batched_r<" +
               std::to_string(r) + ", hsh::JitRun>(d, K, lds); }\n";
    src += "extern \"C\" __global__ __launch_bounds__(1024) void hsh_jit_batched(hsh::DevModel d, int K)\n"
           "{ extern __shared__ double lds[]; hsh::body_batched<hsh::JitRun>(d, K, lds); }\n";
    src += "extern \"C\" __global__ __launch_bounds__(256) void hsh_jit_iso(hsh::DevModel d, int K)\n"
           "{ hsh::body_iso<hsh::JitRun>(d, K); }\n";
    src += "extern \"C\" __global__ __launch_bounds__(256) void hsh_jit_adj_forest(hsh::AdjPack a)\n"
           "{ extern __shared__ double lds[]; hsh::body_adj_forest<hsh::JitAdj>(a, lds); }\n";
    for (int st : {0, 1, 2})
        src += "extern \"C\" __global__ __launch_bounds__(512) void hsh_jit_forest" + std::to_string(st) +
               "(hsh::DevModel d, int K)\n{ extern __shared__ double lds[]; hsh::body_forest<hsh::JitRun, true, " +
               std::to_string(st) + ">(d, K, lds); }\n";
    return src;
}

bool jit_build(const Plan& P, int ncols, JitModule& jm, std::string& err)
{
    auto t0 = std::chrono::steady_clock::now();
    const std::string src = jit_source(P, ncols);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        err = "no device properties";
        return false;
    }
    std::string arch = prop.gcnArchName;
    const std::string key = arch + "\n" + src;
    std::vector<char> code;
    {
        std::lock_guard<std::mutex> lk(g_jit_mu);
        auto it = g_code_cache.find(key);
        if (it != g_code_cache.end()) code = it->second;
    }
    if (code.empty()) {
        hiprtcProgram prog;
        if (hiprtcCreateProgram(&prog, src.c_str(), "hsh_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
            err = "hiprtcCreateProgram failed";
            return false;
        }
        const std::string archopt = "--offload-arch=" + arch.substr(0, arch.find(':'));
        const char* opts[] = {archopt.c_str(), "-O3",      "-ffp-contract=off", "-fno-fast-math",
                              "-std=c++17",    "-Xclang", "-target-feature",   "-Xclang",
                              "-packed-fp32-ops"};
        hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
        if (rc != HIPRTC_SUCCESS) {
            size_t n = 0;
            hiprtcGetProgramLogSize(prog, &n);
            std::string log(n, '\0');
            if (n) hiprtcGetProgramLog(prog, &log[0]);
            err = std::string("hipRTC compile failed: ") + hiprtcGetErrorString(rc) + "\n" + log.substr(0, 2000);
            hiprtcDestroyProgram(&prog);
            return false;
        }
        size_t sz = 0;
        hiprtcGetCodeSize(prog, &sz);
        code.resize(sz);
        hiprtcGetCode(prog, code.data());
        hiprtcDestroyProgram(&prog);
        std::lock_guard<std::mutex> lk(g_jit_mu);
        g_code_cache[key] = code;
    }
    if (hipModuleLoadData(&jm.mod, code.data()) != hipSuccess) {
        err = "hipModuleLoadData failed";
        return false;
    }
    const char* names[] = {"hsh_jit_channel", "hsh_jit_batched_r1", "hsh_jit_batched_r2",
                           "hsh_jit_batched_r4", "hsh_jit_batched", "hsh_jit_iso",
                           "hsh_jit_forest0", "hsh_jit_forest1", "hsh_jit_forest2", "hsh_jit_adj_forest"};
    hipFunction_t* slots[] = {&jm.channel, &jm.batched_r[0], &jm.batched_r[1], &jm.batched_r[2],
                              &jm.batched, &jm.iso, &jm.forest[0], &jm.forest[1], &jm.forest[2],
                              &jm.adj_forest};
    for (int i = 0; i < 10; i++) {
        if (hipModuleGetFunction(slots[i], jm.mod, names[i]) != hipSuccess) {
            err = std::string("hipModuleGetFunction failed: ") + names[i];
            jit_free(jm);
            return false;
        }
    }
    jm.ok = true;
    jm.compile_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

void jit_free(JitModule& jm)
{
    if (jm.mod) (void)hipModuleUnload(jm.mod);
    jm = JitModule();
}

}  // namespace hsh
