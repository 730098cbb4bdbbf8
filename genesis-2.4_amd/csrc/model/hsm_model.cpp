// hsm_model.cpp -- setup compiler for synthetic hsolve populations.
//
// Restates the SETUP/RESET steps of GENESIS's hsolve for plain (asymmetric)
// `compartment` elements carrying tabchannels, spikegens and synchans, and
// emits the arrays the chanmode-4/5 step consumes.  Each step cites the
// reference code it follows.  Compiled -O2 -ffp-contract=off so every double
// expression rounds exactly as the reference's C does.
#include "hsm_model.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

namespace {

// opcodes (genesis/src/hines/hines_defs.h:178-222)
enum {
    COMPT_OP = 100, FCOMPT_OP = 101, LCOMPT_OP = 102, SPIKE_OP = 200,
    CHAN_OP = 3000, CHAN_EK_OP = 3001, ADD_CURR_OP = 3200,
    IPOL1V_OP = 4001, SYN2_OP = 4100, NEWVOLT_OP = 5100,
};
// funcs codes (hines_defs.h:119-128)
enum { FORWARD_ELIM = 0, BACKWARD_ELIM = 0, SET_DIAG = 1, SKIP_DIAG = 2,
       FASTSIBARRAY_ELIM = 3, COPY_ARRAY = 4, SIBARRAY_ELIM = 5, CALC_RESULTS = 6, FINISH = 7 };
// child types used for ordering (hines_defs.h:130-164)
enum { SPIKEGEN_T = 11, TABCHAN_T = 20, SYNCHANC_T = 32 };

const double VTINY = 1.0e-10;     // hines_defs.h:115
const double SING_TINY = 1e-6;    // olf/olf_defs.h:58

thread_local std::string g_err;

struct Child {
    int type;       // SPIKEGEN_T, TABCHAN_T, SYNCHANC_T
    int idx;        // index into the spec's spike/chan/syn arrays
    double Ek;      // for ordering
};

// setup_tab_values (olf/new_interp.c:533-583): fill A and B over xdivs+1
// points with x accumulated by repeated addition of dx.
void setup_tab_values(std::vector<double>& A, std::vector<double>& B,
                      const double* args, int xdivs, double xmin, double xmax,
                      int mode)
{
    A.assign(xdivs + 1, 0.0);
    B.assign(xdivs + 1, 0.0);
    double dx = (xmax - xmin) / xdivs;
    double x = xmin;
    for (int i = 0; i <= xdivs; i++, x += dx) {
        double temp, temp2;
        if (std::fabs(args[4]) < SING_TINY) {
            temp = A[i] = 0.0;
        } else {
            temp2 = args[2] + std::exp((x + args[3]) / args[4]);
            if (std::fabs(temp2) < SING_TINY)
                temp = A[i] = A[i - 1];
            else
                temp = A[i] = (args[0] + args[1] * x) / temp2;
        }
        if (std::fabs(args[9]) < SING_TINY) {
            B[i] = 0.0;
        } else {
            temp2 = args[7] + std::exp((x + args[8]) / args[9]);
            if (std::fabs(temp2) < SING_TINY)
                B[i] = B[i - 1];
            else
                B[i] = (args[5] + args[6] * x) / temp2;
        }
        if (mode == 0) B[i] += temp;       // SETUP_ALPHA: B = alpha + beta
    }
    if (mode == 1) {                       // SETUP_TAU
        for (int i = 0; i <= xdivs; i++) {
            double temp = A[i], temp2 = B[i];
            if (std::fabs(temp) < SING_TINY) {
                A[i] = 0; B[i] = 0;
            } else {
                A[i] = temp2 / temp;
                B[i] = 1.0 / temp;
            }
        }
    }
}

}  // namespace

extern "C" const char* hsm_last_error(void) { return g_err.c_str(); }

// TabInterp, LIN_INTERP branch (olf/new_interp.c:241-252).
extern "C" double hsm_tabinterp(const double* table, int32_t xdivs, double xmin,
                                double invdx, double x)
{
    double lo = (x - xmin) * invdx;
    int ilo = (int)lo;
    if (lo <= 0.0) return table[0];
    if (lo >= xdivs) return table[xdivs];
    double t = table[ilo];
    return (lo - (double)ilo) * (table[ilo + 1] - t) + t;
}

namespace {

template <class T> T* dup(const std::vector<T>& v)
{
    T* p = (T*)std::malloc(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

int fail(const char* msg)
{
    g_err = msg;
    return -1;
}

}  // namespace

// Elimination program of a solver of symmetric compartments: the symmetric
// branch of h_funcs_init (hines_init.c:361-1160, symflag set).  Every
// compartment couples to its parent and to its siblings through the
// junction of their half-resistances; coeff_head(i) = (1 + Ra_i/Ra_parent +
// sum over siblings Ra_i/Ra_s)/2 and coeff_tail(i) = (1 + sum over kids
// Ra_i/Ra_k)/2 (hines_init.c:449-484), and the off-diagonal of row m in
// column n is -dt/(Cm_m Ra_n coeff_m) with the coefficient of the junction
// that joins them.  Eliminating a row's kids and earlier siblings changes
// the off-diagonals to its parent and later siblings: the program copies
// those into working ravals slots at the start of the row (COPY_ARRAY),
// updates them as each pivot row is eliminated (FASTSIBARRAY_ELIM,
// SIBARRAY_ELIM), and FORWARD_ELIM / SIBARRAY_ELIM read the working slots
// where the ravals stream reaches them.  A working slot is a placeholder
// (-1) in funcs until the ravals position that the stream reads it from is
// known; funcs entries holding -(placeholder index) are copies resolved at
// the same time.  The slot bookkeeping follows the reference's ravptr table
// entry for entry (per row: [0] the right off-diagonal to the parent,
// [1..maxkids] the kids' left off-diagonals, [maxkids+1..2 maxkids] the
// siblings' off-diagonals; -1 unset, 0 unused, > 0 a placeholder's funcs
// index, < -1 minus a ravals index), so the program, its ravals and the
// diagonal terms are the reference's for the same tree.  `prod` is a float
// product when element fields are float (two float fields multiplied
// without a double operand, as at hines_init.c:632, :718, :887).
namespace {
struct SymOut {
    std::vector<int> funcs;
    std::vector<double> ravals, diag;
};

int sym_funcs(int n, double cdt, const std::vector<int>& elmnum, const std::vector<int>& hnum,
              const std::vector<int>& parents, const std::vector<int>& nkids, const std::vector<int>& kstart,
              const std::vector<int>& kids, const std::vector<double>& Cmv, const std::vector<double>& Rmv,
              const std::vector<double>& Rav, bool ff, SymOut& out)
{
    auto prod = [ff](double a, double b) { return ff ? (double)((float)a * (float)b) : a * b; };
    auto kid = [&](int e, int j) { return kids[kstart[e] + j]; };
    auto bug = [](int k) {
        g_err = "hsm_build: symmetric elimination program inconsistent (check " + std::to_string(k) + ")";
        return -1;
    };
    int maxkids = 1;
    for (int i = 0; i < n; i++) maxkids = std::max(maxkids, nkids[i]);
    const int W2 = 2 * maxkids + 1;
    std::vector<int> ravptr((size_t)n * W2, -1);
    auto RP = [&](int row, int l) -> int& { return ravptr[(size_t)row * W2 + l]; };
    std::vector<double> coeffs(2 * (size_t)n);
    for (int i = 0; i < n; i++) {
        const int ci = elmnum[i];
        const double Ra = Rav[ci];
        double temp = 0.0;
        const int pn = parents[ci];
        if (pn >= 0) {
            temp = Ra / Rav[pn];
            for (int j = 0; j < nkids[pn]; j++) {
                const int k = kid(pn, j);
                if (k != ci) temp += Ra / Rav[k];
            }
        }
        coeffs[2 * i] = (1.0 + temp) / 2.0;
        temp = 0.0;
        for (int j = 0; j < nkids[ci]; j++) temp += Ra / Rav[kid(ci, j)];
        coeffs[2 * i + 1] = (1.0 + temp) / 2.0;
    }
    std::vector<int>& f = out.funcs;
    std::vector<double>& ra = out.ravals;
    std::vector<double>& dg = out.diag;
    f.clear();
    ra.clear();
    dg.assign(n, 0.0);
    const double dt = cdt;
    // a placeholder at funcs index m takes ravals index q, with its copies
    auto resolve = [&](int m, int q, bool need_copy) -> bool {
        f[m] = q;
        bool found = false;
        for (size_t l = (size_t)m; l < f.size(); l++)
            if (f[l] == -m) {
                f[l] = q;
                found = true;
            }
        return found || !need_copy;
    };
    int skipdiag = 0, self = 0;
    for (int i = 0; i < n; i++) {
        const int ci = elmnum[i];
        const double Cm = Cmv[ci];
        dg[i] = 1.0 + dt / (Cm * Rmv[ci]);
        const int pn = parents[ci];
        const int nkid = nkids[ci];
        for (int l = nkid + 1; l <= maxkids; l++) RP(i, l) = 0;
        int hassiblefts = 0;
        if (pn > -1) {
            const int np = nkids[pn];
            int j = 0;
            for (; j < np; j++)
                if (kid(pn, j) == ci) {
                    self = j;
                    break;
                }
            if (np && j == np) return bug(0);
            if (np > 1) {
                if (self > 0) {
                    f.push_back(COPY_ARRAY);
                    RP(i, 0) = (int)f.size();
                    f.push_back(-1);
                    const double temp = -dt / (Cm * Rav[pn] * coeffs[2 * i]);
                    ra.push_back(temp);
                    dg[i] -= temp;
                    hassiblefts = 1;
                }
            } else {
                RP(i, 0) = 0;
            }
            if (i == 0) {
                RP(0, maxkids + 1) = 0;
                for (int l = maxkids + np + 1; l <= 2 * maxkids; l++) RP(0, l) = 0;
                continue;
            }
        } else {
            RP(i, 0) = 0;
        }
        if (skipdiag) {
            skipdiag = 0;
        } else if (nkid || hassiblefts || pn == -1) {
            if (i > 1) f.push_back(SET_DIAG);
        } else if (i > 1) {
            f.push_back(SKIP_DIAG);
            skipdiag = 1;
        }
        for (int j = 0; j < nkid; j++) {
            const int k = kid(ci, j);
            const int row = hnum[k];
            f.push_back(FORWARD_ELIM);
            f.push_back(2 * row);
            if (j == 0) {
                const double temp = -dt / (Cm * Rav[k] * coeffs[2 * i + 1]);
                ra.push_back(temp);
                dg[i] -= temp;
            } else {
                const int m = RP(i, j + 1);
                if (m <= 0 || f[m] != -1) return bug(1);
                if (!resolve(m, (int)ra.size(), true)) return bug(-1);
                ra.push_back(0.0);    // a working slot: COPY_ARRAY fills it each step
            }
            RP(i, j + 1) = 0;
            if (j == 0) {
                const int m = RP(row, 0);
                if (m < -1) {
                    ra.push_back(ra[-m]);
                } else {
                    const double temp = -dt / (prod(Cmv[k], Rav[ci]) * coeffs[2 * row]);
                    ra.push_back(temp);
                    dg[row] -= temp;
                }
            } else {
                const int m = RP(row, 0);
                if (m <= 0 || f[m] != -1) return bug(2);
                resolve(m, (int)ra.size(), false);
                ra.push_back(0.0);
            }
            RP(row, 0) = -(int)(ra.size() - 1);
            if (j == 0) {
                for (int l = 1; l < nkid; l++) {
                    f.push_back(COPY_ARRAY);
                    RP(i, l + 1) = (int)f.size();
                    f.push_back(-1);
                    const int k2 = kid(ci, l);
                    const double temp = -dt / (Cm * Rav[k2] * coeffs[2 * i + 1]);
                    ra.push_back(temp);
                    dg[i] -= temp;
                }
            }
            for (int l = j + 1; l < nkid; l++) {
                f.push_back(SIBARRAY_ELIM);
                int m = RP(row, maxkids + l + 1);
                f.push_back(-m);
                if (m == 0) return bug(3);
                m = RP(i, l + 1);
                if (m <= 0) return bug(4);
                f.push_back(-m);
            }
        }
        int nsib = 0;   // sibling off-diagonals of this row eliminated so far
        if (pn > -1) {
            const int np = nkids[pn];
            RP(i, maxkids + 1) = 0;
            RP(i, maxkids + 1 + self) = 0;
            for (int l = maxkids + np + 1; l <= 2 * maxkids; l++) RP(i, l) = 0;
            if (np < 2) continue;
            for (int j = 0; j < np; j++) {
                const int k = kid(pn, j);
                const int row = hnum[k];
                if (np > 2 && self > 0 && j > 0) {
                    const double temp = -dt / (Cm * Rav[k] * coeffs[2 * i]);
                    ra.push_back(temp);
                    dg[i] -= temp;
                    f.push_back(COPY_ARRAY);
                    RP(i, maxkids + j + 1) = (int)f.size();
                    f.push_back(-1);
                }
                if (row >= i) continue;
                f.push_back(FORWARD_ELIM);
                f.push_back(2 * row);
                if (nsib == 0) {
                    const double temp = -dt / (Cm * Rav[k] * coeffs[2 * i]);
                    ra.push_back(temp);
                    dg[i] -= temp;
                } else {
                    const int m = RP(i, maxkids + j + 1);
                    if (m <= 0 || f[m] != -1) return bug(5);
                    if (!resolve(m, (int)ra.size(), true)) return bug(-5);
                    RP(i, maxkids + j + 1) = (int)ra.size();
                    ra.push_back(0.0);
                }
                if (nsib == 0) {
                    const double temp = -dt / (prod(Cmv[k], Rav[ci]) * coeffs[2 * row]);
                    RP(row, maxkids + self + 1) = -(int)ra.size();
                    ra.push_back(temp);
                    dg[row] -= temp;
                } else {
                    const int m = RP(row, maxkids + j + 1);
                    if (m <= 0 || f[m] != -1) return bug(6);
                    if (!resolve(m, (int)ra.size(), true)) return bug(-6);
                    ra.push_back(0.0);
                    RP(row, maxkids + j + 1) = 0;
                }
                if (self > 0) {
                    int m = RP(row, 0);
                    if (m == -1) {
                        f.push_back(FASTSIBARRAY_ELIM);
                        RP(row, 0) = -(int)ra.size();
                        const double temp = -dt / (prod(Cmv[k], Rav[pn]) * coeffs[2 * row]);
                        ra.push_back(temp);
                        dg[row] -= temp;
                        m = RP(i, 0);
                        if (m <= 0) return bug(-7);
                        f.push_back(-m);
                    } else {
                        f.push_back(SIBARRAY_ELIM);
                        if (m == -1 || m == 0 || (j > 0 && m < 0)) return bug(7);
                        f.push_back(-m);
                        m = RP(i, 0);
                        if (m <= 0) return bug(-7);
                        f.push_back(-m);
                    }
                    for (int l = j + 1; l < np; l++) {
                        if (l == self) continue;
                        f.push_back(SIBARRAY_ELIM);
                        m = RP(row, maxkids + l + 1);
                        if (m == -1 || m == 0 || (j > 0 && m < 0)) return bug(8);
                        f.push_back(-m);
                        m = RP(i, maxkids + l + 1);
                        if (m <= 0) return bug(-8);
                        f.push_back(-m);
                    }
                    nsib++;
                }
            }
        } else {
            for (int l = maxkids + 1; l <= 2 * maxkids; l++) RP(i, l) = 0;
        }
    }
    for (int v : f)
        if (v < 0) return bug(9);
    f.push_back(FINISH);
    // backward pass: rows from the second last down; the parent's
    // off-diagonal first, then those to later siblings (hines_init.c:1020-1160)
    for (int i = n - 2; i >= 0; i--) {
        const int ci = elmnum[i];
        const double Cm = Cmv[ci];
        const int pn = parents[ci];
        if (pn != -1) {
            const int np = nkids[pn];
            for (int j = 0; j < np; j++)
                if (kid(pn, j) == ci) {
                    self = j;
                    break;
                }
            const int row = hnum[pn];
            if (self > 0) {
                f.push_back(SIBARRAY_ELIM);
                if (RP(i, 0) >= 0) return bug(11);
                f.push_back(-RP(i, 0));
                f.push_back(2 * row);
            } else {
                f.push_back(BACKWARD_ELIM);
                f.push_back(2 * row);
                ra.push_back(-dt / (Cm * Rav[pn] * coeffs[2 * i]));
            }
            for (int j = 0; j < np; j++) {
                const int k = kid(pn, j);
                const int rk = hnum[k];
                if (rk <= i) continue;
                if (j > 0) {
                    f.push_back(SIBARRAY_ELIM);
                    if (RP(i, maxkids + j + 1) >= 0) return bug(12);
                    f.push_back(-RP(i, maxkids + j + 1));
                    f.push_back(2 * rk);
                } else {
                    f.push_back(BACKWARD_ELIM);
                    f.push_back(2 * rk);
                    ra.push_back(-dt / (Cm * Rav[k] * coeffs[2 * i]));
                }
            }
        }
        f.push_back(CALC_RESULTS);
    }
    f.push_back(FINISH);
    return 0;
}
}  // namespace

extern "C" int32_t hsm_build(const hsm_spec* s, hsm_out** out_p)
{
    *out_p = nullptr;
    const int n = s->ncomp;
    if (n <= 0) return fail("hsm_build: empty compartment list");
    if (s->chanmode != 4 && s->chanmode != 5) return fail("hsm_build: chanmode must be 4 or 5");
    const bool crank = (s->method == HSM_CRANK);
    // Element fields: GENESIS declares Em, Rm, Cm, Ra, inject, initVm, Ek and
    // Gbar as C float (segment/seg_struct.h:209-221, sim/struct_defs.h:118-122,
    // olf/olf_struct.h:142-155).  With float_fields the values are rounded to
    // float, and the two setup expressions that combine two float fields
    // without a double operand are evaluated in float, as x86-64 C does:
    // chip Em/Rm (hines_data.c:112) and the kid coefficient
    // -dt/(link->Cm*link->Ra) (hines_init.c:733).
    const bool ff = s->float_fields != 0;
    auto F = [ff](double x) { return ff ? (double)(float)x : x; };
    std::vector<double> fCm(n), fRm(n), fRa(n), fEm(n), fV0(n), fInj(n);
    for (int i = 0; i < n; i++) {
        fCm[i] = F(s->Cm[i]); fRm[i] = F(s->Rm[i]); fRa[i] = F(s->Ra[i]);
        fEm[i] = F(s->Em[i]); fV0[i] = F(s->initVm[i]); fInj[i] = F(s->inject[i]);
    }
    auto em_over_rm = [&](int e) {
        return ff ? (double)((float)fEm[e] / (float)fRm[e]) : fEm[e] / fRm[e];
    };
    auto cm_ra = [&](int e) {
        return ff ? (double)((float)fCm[e] * (float)fRa[e]) : fCm[e] * fRa[e];
    };
    // Coefficient dt is halved for Crank-Nicolson (hines_init.c:396-399,
    // hines_data.c:91-94).
    const double cdt = crank ? s->dt / 2.0 : s->dt;

    // ---- topology: parents, kids, Hines numbering (hines_init.c:40-345)
    std::vector<int> parents(n), nkids(n, 0);
    for (int i = 0; i < n; i++) {
        int p = s->parent[i];
        if (p < -1 || p >= n || p == i) return fail("hsm_build: bad parent index");
        parents[i] = p;
        if (p >= 0) nkids[p]++;
    }
    // kids filled in element order (hines_init.c:252-269)
    std::vector<int> kstart(n + 1, 0);
    for (int i = 0; i < n; i++) kstart[i + 1] = kstart[i] + nkids[i];
    std::vector<int> kids(std::max(1, kstart[n])), kcur(n, 0);
    for (int j = 0; j < n; j++)
        if (parents[j] != -1) { int k = parents[j]; kids[kstart[k] + kcur[k]++] = j; }

    // do_hnum (hines_init.c:317-345), iterative form of the same recursion:
    // number the element, then its leaf kids, then recurse into non-leaf
    // kids in order.
    std::vector<int> hnum(n, -1), elmnum(n, -1);
    int hcount = n - 1;
    {
        struct Frame { int c; int next; };
        std::vector<Frame> st;
        for (int r = 0; r < n; r++) {
            if (parents[r] != -1) continue;
            // visit r
            hnum[r] = hcount; elmnum[hcount] = r; hcount--;
            for (int q = kstart[r]; q < kstart[r + 1]; q++) {
                int kid = kids[q];
                if (nkids[kid] == 0) { hnum[kid] = hcount; elmnum[hcount] = kid; hcount--; }
            }
            st.push_back({r, kstart[r]});
            while (!st.empty()) {
                Frame& f = st.back();
                int found = -1;
                while (f.next < kstart[f.c + 1]) {
                    int kid = kids[f.next++];
                    if (nkids[kid] > 0) { found = kid; break; }
                }
                if (found < 0) { st.pop_back(); continue; }
                int c = found;
                hnum[c] = hcount; elmnum[hcount] = c; hcount--;
                for (int q = kstart[c]; q < kstart[c + 1]; q++) {
                    int kid = kids[q];
                    if (nkids[kid] == 0) { hnum[kid] = hcount; elmnum[hcount] = kid; hcount--; }
                }
                st.push_back({c, kstart[c]});
            }
        }
    }
    if (hcount != -1) return fail("hsm_build: cyclic compartment graph");
    // sort kids by Hines number (hines_init.c:271-295)
    for (int i = 0; i < n; i++)
        std::sort(kids.begin() + kstart[i], kids.begin() + kstart[i + 1],
                  [&](int a, int b) { return hnum[a] < hnum[b]; });

    // ---- children of each compartment, in solver order
    // hchild_compare (hines_child.c:1990-2083) orders spikegens, then
    // tabchannels by ascending Ek, then synchans; the insertion sort
    // (hines_child.c:1262-1277) places a child before earlier children it does
    // not follow, so ties end up in reverse creation order.
    std::vector<std::vector<Child>> ch(n);
    for (int i = 0; i < s->nspike; i++) {
        int c = s->spike_comp[i];
        if (c < 0 || c >= n) return fail("hsm_build: bad spikegen compartment");
        ch[c].push_back({SPIKEGEN_T, i, 0.0});
    }
    for (int i = 0; i < s->nchan; i++) {
        int c = s->chan_comp[i];
        int t = s->chan_type[i];
        if (c < 0 || c >= n || t < 0 || t >= s->nchantypes) return fail("hsm_build: bad channel record");
        ch[c].push_back({TABCHAN_T, i, F(s->chantypes[t].Ek)});
    }
    for (int i = 0; i < s->nsyn; i++) {
        int c = s->syn_comp[i];
        int t = s->syn_type[i];
        if (c < 0 || c >= n || t < 0 || t >= s->nsyntypes) return fail("hsm_build: bad synchan record");
        ch[c].push_back({SYNCHANC_T, i, F(s->syntypes[t].Ek)});
    }
    auto before = [](const Child& a, const Child& b) {
        if (a.type < b.type) return true;
        if (a.type == b.type && a.type == TABCHAN_T) return a.Ek < b.Ek;
        return false;
    };
    int nchildren = 0;
    for (int c = 0; c < n; c++) {
        std::vector<Child> v = ch[c];          // creation order per type
        // spikegens, then tabchannels, then synchans were pushed type-major;
        // restore creation order within the compartment is type-major
        // already, which is what the comparator's first key produces.
        std::vector<Child> o;
        for (const Child& x : v) {
            size_t m = o.size();
            o.push_back(x);
            while (m > 0 && !before(o[m - 1], x)) { o[m] = o[m - 1]; m--; }
            o[m] = x;
        }
        ch[c] = o;
        nchildren += (int)o.size();
    }

    // ---- tables: register gate tables in first-use order, build composite
    // table (hines_chip_init.c:148-313, :1262-1345).
    const int xdivs = s->xdivs;
    const double xmin = s->xmin;
    double tdx = (s->xmax - s->xmin) / (double)xdivs;    // create_interpol
    double invdx = (std::fabs(tdx) > 1e-50) ? 1.0 / tdx : 1.0;
    // table column of gate g of channel type t, -1 = not yet registered
    std::vector<int> gatecol(3 * std::max(1, s->nchantypes), -1);
    std::vector<std::vector<double>> rawA(3 * std::max(1, s->nchantypes)), rawB(3 * std::max(1, s->nchantypes));
    std::vector<int> colorder;   // gate ids in column order
    for (int t = 0; t < s->nchantypes; t++)
        for (int g = 0; g < 3; g++) {
            const hsm_gate_spec& gs = s->chantypes[t].gate[g];
            if (gs.power == 0) continue;
            if (gs.power < 1 || gs.power > 6)
                return fail("hsm_build: gate power must be in 1..6 (hines_chip_init.c:330)");
            setup_tab_values(rawA[3 * t + g], rawB[3 * t + g], gs.args, xdivs, s->xmin, s->xmax, gs.mode);
        }
    std::vector<char> needvolt(n, 0);
    for (int r = 0; r < n; r++) {
        int e = elmnum[r];
        for (const Child& c : ch[e]) {
            if (c.type != TABCHAN_T) continue;
            int t = s->chan_type[c.idx];
            for (int g = 0; g < 3; g++) {
                if (s->chantypes[t].gate[g].power == 0) continue;
                needvolt[r] = 1;
                if (gatecol[3 * t + g] < 0) {
                    gatecol[3 * t + g] = 2 * (int)colorder.size();
                    colorder.push_back(3 * t + g);
                }
            }
        }
    }
    const int ncols = 2 * (int)colorder.size();
    const double tby2 = s->dt / 2.0;
    std::vector<double> tablist, xvals;
    if (ncols > 0) {
        tablist.assign((size_t)(xdivs + 2) * ncols, 0.0);
        xvals.assign(xdivs + 2, 0.0);
        double x = xmin, dx = 1.0 / invdx;
        for (int j = 0; j < xdivs + 2; j++) { xvals[j] = x; x += dx; }
        for (size_t q = 0; q < colorder.size(); q++) {
            int id = colorder[q];
            int t = id / 3, g = id % 3;
            const hsm_gate_spec& gs = s->chantypes[t].gate[g];
            int l = 2 * (int)q;
            for (int j = 0; j <= xdivs; j++) {
                if (!gs.instant) {
                    tablist[(size_t)j * ncols + l] = 1.0 + tby2 * rawB[id][j];
                    tablist[(size_t)j * ncols + l + 1] = s->dt * rawA[id][j];
                } else {
                    tablist[(size_t)j * ncols + l] = rawB[id][j];
                    tablist[(size_t)j * ncols + l + 1] = rawA[id][j];
                }
            }
        }
        // guard row: the reference reads one row past the table when Vm is at
        // or above xmax (hines_chip.c:181-186 with filo == xdivs); with
        // vipol == 0 the interpolant is then the top row, which a copy of the
        // top row reproduces.
        for (int l = 0; l < ncols; l++)
            tablist[(size_t)(xdivs + 1) * ncols + l] = tablist[(size_t)xdivs * ncols + l];
    }

    // ---- synapse table (hines_chip_init.c:425-463; newconn/synchan.c:535-559)
    std::vector<double> stab;
    std::vector<int> syn_stab(std::max(1, s->nsyn), 0);
    for (int i = 0; i < s->nsyn; i++) {
        const hsm_syntype& st = s->syntypes[s->syn_type[i]];
        double dt = s->dt;
        double xconst1 = st.tau1 * (1.0 - std::exp(-dt / st.tau1));
        double xconst2 = std::exp(-dt / st.tau1);
        double yconst1 = st.tau2 * (1.0 - std::exp(-dt / st.tau2));
        double yconst2 = std::exp(-dt / st.tau2);
        double e3 = xconst1 / dt;      // SYNCHANC_T: normalised by dt
        int found = -1;
        for (size_t j = 0; j < stab.size(); j += 4) {
            if (std::fabs(stab[j] - xconst2) < VTINY && std::fabs(stab[j + 1] - yconst1) < VTINY &&
                std::fabs(stab[j + 2] - yconst2) < VTINY &&
                (std::fabs(stab[j + 3] - xconst1) < VTINY || std::fabs(stab[j + 3] - e3) < VTINY)) {
                found = (int)j; break;
            }
        }
        if (found < 0) {
            found = (int)stab.size();
            stab.push_back(xconst2); stab.push_back(yconst1);
            stab.push_back(yconst2); stab.push_back(e3);
        }
        syn_stab[i] = found;
    }

    // ---- channel program and chip layout (hines_chip_init.c:762-1243)
    std::vector<int> ops;
    std::vector<double> chip;
    std::vector<int> compops(n), compchips(n);
    std::vector<int> spike_row(std::max(1, s->nspike)), spike_op(std::max(1, s->nspike));
    std::vector<int> syn_op(std::max(1, s->nsyn)), syn_chip(std::max(1, s->nsyn));
    std::vector<double> givals(1, 0.0);   // givals[0] is never used
    std::vector<int> compgiv(n);
    std::vector<double> vm(n);
    double lastEk = 1.0e30;
    int childno = 0;
    for (int r = 0; r < n; r++) {
        int e = elmnum[r];
        compops[r] = (int)ops.size();
        compchips[r] = (int)chip.size();
        ops.push_back(r == 0 ? FCOMPT_OP : COMPT_OP);
        // RESET values of the compartment (hines_data.c:80-125)
        chip.push_back(em_over_rm(e));
        chip.push_back(fInj[e]);
        vm[r] = fV0[e];
        givals.push_back(fRm[e]);        // Rm, to compute the leak current
        givals.push_back(0.0);             // leak Im
        if (needvolt[r]) ops.push_back(NEWVOLT_OP);
        for (const Child& c : ch[e]) {
            if (c.type == SPIKEGEN_T) {
                spike_row[c.idx] = r;
                spike_op[c.idx] = (int)ops.size();
                ops.push_back(SPIKE_OP);
                ops.push_back(0);          // counter (chip_put_spike: cleared)
                ops.push_back((int)((F(s->spike_refract[c.idx]) / s->dt) + 1.9));
                chip.push_back(F(s->spike_thresh[c.idx]));
                childno++;
                continue;
            }
            double Ek = c.Ek;
            double gmax;
            if (c.type == TABCHAN_T) {
                gmax = F(s->chan_gbar[c.idx]);
            } else {
                // synchan conductance scale `norm` (newconn/synchan.c:545-556),
                // stored where a tabchannel keeps Gbar (hines_data.c:504)
                const hsm_syntype& st = s->syntypes[s->syn_type[c.idx]];
                double g = s->syn_gmax[c.idx];
                if (st.tau1 == st.tau2) {
                    gmax = g * 2.718281828459045 / st.tau1;
                } else {
                    double tpeak = st.tau1 * st.tau2 * std::log(st.tau1 / st.tau2) / (st.tau1 - st.tau2);
                    gmax = g * (st.tau1 - st.tau2) /
                           (st.tau1 * st.tau2 * (std::exp(-tpeak / st.tau1) - std::exp(-tpeak / st.tau2)));
                }
            }
            // Ek source (hines_chip_init.c:1029-1049)
            if (std::fabs(Ek - lastEk) < VTINY) {
                ops.push_back(CHAN_OP);
                chip.push_back(gmax);
            } else {
                ops.push_back(CHAN_EK_OP);
                chip.push_back(Ek);
                chip.push_back(gmax);
                lastEk = Ek;
            }
            if (c.type == TABCHAN_T) {
                int t = s->chan_type[c.idx];
                for (int g = 0; g < 3; g++) {
                    const hsm_gate_spec& gs = s->chantypes[t].gate[g];
                    if (gs.power == 0) continue;
                    int col = gatecol[3 * t + g];
                    ops.push_back(IPOL1V_OP);
                    ops.push_back(col);
                    ops.push_back(gs.instant ? -gs.power : gs.power);
                    // RESET: X = A(V)/B(V) on the element's own tables
                    // (olf/tabchannel.c:274-285)
                    double v = fV0[e];
                    double A = hsm_tabinterp(rawA[3 * t + g].data(), xdivs, xmin, invdx, v);
                    double B = hsm_tabinterp(rawB[3 * t + g].data(), xdivs, xmin, invdx, v);
                    chip.push_back(A / B);
                }
                ops.push_back(ADD_CURR_OP);
            } else {
                syn_op[c.idx] = (int)ops.size();
                ops.push_back(SYN2_OP);
                ops.push_back(syn_stab[c.idx]);
                ops.push_back(-1);         // no pending event
                ops.push_back(childno);
                syn_chip[c.idx] = (int)chip.size();
                chip.push_back(0.0);       // X
                chip.push_back(0.0);       // Y
            }
            givals.push_back(0.0); givals.push_back(0.0); givals.push_back(0.0);
            childno++;
        }
        chip.push_back(cdt / fCm[e]);    // tbyc
        chip.push_back(0.0);               // diagterm, filled below
        givals.push_back(0.0);             // total Im
        compgiv[r] = (int)givals.size() - 1;
    }
    ops.push_back(LCOMPT_OP);

    // ---- elimination program (hines_init.c:361-1273)
    std::vector<int> funcs;
    std::vector<double> ravals;
    std::vector<double> diagterms(n);
    int skipdiag = 0;
    if (s->symmetric) {
        SymOut so;
        if (sym_funcs(n, cdt, elmnum, hnum, parents, nkids, kstart, kids, fCm, fRm, fRa, ff, so) != 0) return -1;
        funcs = so.funcs;
        ravals = so.ravals;
        diagterms = so.diag;
    } else {
        for (int i = 0; i < n; i++) {
            int ci = elmnum[i];
            double Cm = fCm[ci];
            diagterms[i] = 1.0 + cdt / (Cm * fRm[ci]);
            int parentno = parents[ci];
            int nk = nkids[ci];
            if (skipdiag) {
                skipdiag = 0;
            } else if (nk || parentno == -1) {
                if (i > 1) funcs.push_back(SET_DIAG);
            } else {
                if (i > 1) { funcs.push_back(SKIP_DIAG); skipdiag = 1; }
            }
            for (int q = kstart[ci]; q < kstart[ci + 1]; q++) {
                int k = kids[q];
                int row = hnum[k];
                funcs.push_back(FORWARD_ELIM);
                funcs.push_back(2 * row);
                double temp = -cdt / (Cm * fRa[k]);
                ravals.push_back(temp);
                diagterms[i] -= temp;
                temp = -cdt / cm_ra(k);
                ravals.push_back(temp);
                diagterms[row] -= temp;
            }
        }
        funcs.push_back(FINISH);
        for (int i = n - 2; i >= 0; i--) {
            int ci = elmnum[i];
            int parentno = parents[ci];
            if (parentno != -1) {
                funcs.push_back(BACKWARD_ELIM);
                funcs.push_back(2 * hnum[parentno]);
                ravals.push_back(-cdt / (fCm[ci] * fRa[ci]));
            }
            funcs.push_back(CALC_RESULTS);
        }
        funcs.push_back(FINISH);
    }
    // diagterms into the last chip slot of each compartment
    // (hines_init.c:1214-1229)
    for (int i = 0; i < n; i++) {
        int j = (i == n - 1) ? (int)chip.size() - 1 : compchips[i + 1] - 1;
        chip[j] = diagterms[i];
    }

    std::vector<int> parent_row(n);
    for (int r = 0; r < n; r++) {
        int p = parents[elmnum[r]];
        parent_row[r] = (p < 0) ? -1 : hnum[p];
    }

    hsm_out* o = (hsm_out*)std::calloc(1, sizeof(hsm_out));
    o->ncompts = n;
    o->nchips = (int)chip.size();
    o->nops = (int)ops.size();
    o->nfuncs = (int)funcs.size();
    o->nravals = (int)ravals.size();
    o->ncols = ncols;
    o->xdivs = (ncols > 0) ? xdivs : 0;
    o->ngivals = (int)givals.size();
    o->nstab = (int)stab.size() / 4;
    o->method = s->method;
    o->chanmode = s->chanmode;
    o->nchildren = nchildren;
    o->xmin = (ncols > 0) ? xmin : 0.0;
    o->invdx = (ncols > 0) ? invdx : 0.0;
    o->dt = s->dt;
    o->ops = dup(ops);
    o->chip = dup(chip);
    o->tablist = dup(tablist);
    o->xvals = dup(xvals);
    o->funcs = dup(funcs);
    o->ravals = dup(ravals);
    o->vm = dup(vm);
    o->givals = dup(givals);
    o->stablist = dup(stab);
    o->compops = dup(compops);
    o->compchips = dup(compchips);
    o->elmnum = dup(elmnum);
    o->hnum = dup(hnum);
    o->parent_row = dup(parent_row);
    spike_row.resize(s->nspike); spike_op.resize(s->nspike);
    syn_op.resize(s->nsyn); syn_chip.resize(s->nsyn);
    o->spike_row = dup(spike_row);
    o->spike_op = dup(spike_op);
    o->syn_op = dup(syn_op);
    o->syn_chip = dup(syn_chip);
    *out_p = o;
    return 0;
}

extern "C" void hsm_free(hsm_out* o)
{
    if (!o) return;
    void* ptrs[] = {o->ops, o->chip, o->tablist, o->xvals, o->funcs, o->ravals, o->vm,
                    o->givals, o->stablist, o->compops, o->compchips, o->elmnum, o->hnum,
                    o->parent_row, o->spike_row, o->spike_op, o->syn_op, o->syn_chip};
    for (void* p : ptrs) std::free(p);
    std::free(o);
}

// ---------------------------------------------------------------------------
// Synthetic populations

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
};

// HH squid-axon gates as set up by setupalpha in
// genesis/Scripts/benchmark/hh1952_ap_verify.g:49-62, parameterised by the
// table rest potential E.
void hh_types(double E, double ENa, double EK, hsm_chantype* na, hsm_chantype* k)
{
    std::memset(na, 0, sizeof(*na));
    std::memset(k, 0, sizeof(*k));
    na->Ek = ENa;
    hsm_gate_spec& nx = na->gate[0];
    nx.power = 3;
    double nxa[10] = {100000.0 * (0.025 + E), -100000.0, -1.0, -1.0 * (0.025 + E), -0.01,
                      4000.0, 0.0, 0.0, 0.0 - E, 0.018};
    std::memcpy(nx.args, nxa, sizeof(nxa));
    hsm_gate_spec& ny = na->gate[1];
    ny.power = 1;
    double nya[10] = {70.0, 0.0, 0.0, 0.0 - E, 0.02, 1000.0, 0.0, 1.0, -0.03 - E, -0.01};
    std::memcpy(ny.args, nya, sizeof(nya));
    k->Ek = EK;
    hsm_gate_spec& kx = k->gate[0];
    kx.power = 4;
    double kxa[10] = {10000.0 * (0.010 + E), -10000.0, -1.0, -1.0 * (0.010 + E), -0.01,
                      125.0, 0.0, 0.0, 0.0 - E, 0.08};
    std::memcpy(kx.args, kxa, sizeof(kxa));
}

}  // namespace

extern "C" int32_t hsm_build_population(const hsm_pop* p, hsm_out** out)
{
    *out = nullptr;
    if (p->n_neurons <= 0 || p->nsizes <= 0) return fail("hsm_build_population: empty population");
    // one generator per neuron, seeded from (seed, global neuron index), so
    // any contiguous sub-range [first_neuron, first_neuron+n_neurons) of a
    // population is built identically on its own
    SplitMix64 seeder(p->seed);
    const uint64_t salt = seeder.next();
    const double PI = 3.141592653589793;
    std::vector<double> cum(p->nsizes);
    double tot = 0;
    for (int i = 0; i < p->nsizes; i++) { tot += p->probs[i]; cum[i] = tot; }
    std::vector<int> parent;
    std::vector<double> Cm, Rm, Ra, Em, initVm, inject;
    std::vector<int> chan_comp, chan_type;
    std::vector<double> chan_gbar;
    const int max_kids = p->max_kids > 0 ? p->max_kids : 3;
    for (int nn = 0; nn < p->n_neurons; nn++) {
        SplitMix64 rng(salt ^ (0x9E3779B97F4A7C15ull * (uint64_t)(p->first_neuron + nn + 1)));
        rng.next();
        int size = p->sizes[0];
        if (p->nsizes > 1) {
            double u = rng.uniform() * tot;
            int q = 0;
            while (q < p->nsizes - 1 && u >= cum[q]) q++;
            size = p->sizes[q];
        }
        int base = (int)parent.size();
        std::vector<int> nk(size, 0);
        for (int c = 0; c < size; c++) {
            double len, dia;
            if (c == 0) {
                parent.push_back(-1);
                len = p->soma_len; dia = p->soma_dia;
            } else {
                int par;
                do { par = (int)(rng.uniform() * c); } while (nk[par] >= max_kids);
                nk[par]++;
                parent.push_back(base + par);
                len = rng.uniform(p->dend_len_lo, p->dend_len_hi);
                dia = rng.uniform(p->dend_dia_lo, p->dend_dia_hi);
            }
            double area = PI * dia * len;
            double xarea = PI * (dia / 2) * (dia / 2);
            Cm.push_back(p->CM * area);
            Rm.push_back(p->RM / area);
            Ra.push_back(p->RA * len / xarea);
            Em.push_back(p->Em);
            initVm.push_back(p->Em);
            if (c == 0) {
                inject.push_back(rng.uniform(p->inject_lo, p->inject_hi));
                if (p->gNa > 0) { chan_comp.push_back(base); chan_type.push_back(0); chan_gbar.push_back(p->gNa * area); }
                if (p->gK > 0) { chan_comp.push_back(base); chan_type.push_back(1); chan_gbar.push_back(p->gK * area); }
            } else {
                inject.push_back(0.0);
                if (p->gKdr > 0) { chan_comp.push_back(base + c); chan_type.push_back(1); chan_gbar.push_back(p->gKdr * area); }
            }
        }
    }
    hsm_chantype types[2];
    hh_types(p->erest_act, p->erest_act + 0.115, p->erest_act - 0.012, &types[0], &types[1]);
    hsm_spec s;
    std::memset(&s, 0, sizeof(s));
    s.dt = p->dt;
    s.method = p->method ? p->method : HSM_BEULER;
    s.chanmode = 4;
    s.xdivs = 3000;
    s.float_fields = p->float_fields;
    s.xmin = -0.1;
    s.xmax = 0.05;
    s.ncomp = (int)parent.size();
    s.parent = parent.data();
    s.Cm = Cm.data(); s.Rm = Rm.data(); s.Ra = Ra.data();
    s.Em = Em.data(); s.initVm = initVm.data(); s.inject = inject.data();
    s.nchantypes = 2;
    s.chantypes = types;
    s.nchan = (int)chan_comp.size();
    s.chan_comp = chan_comp.data();
    s.chan_type = chan_type.data();
    s.chan_gbar = chan_gbar.data();
    return hsm_build(&s, out);
}
