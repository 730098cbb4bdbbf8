// hsh_api.cpp -- C ABI of libhsolve_hip (declared in include/hsolve_hip.h).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
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

template <class T> hipError_t dalloc(T** p, size_t n)
{
    return hipMalloc((void**)p, std::max<size_t>(1, n) * sizeof(T));
}

constexpr int SPIKE_CAP = 1 << 20;
constexpr int GRAPH_MAX_LAUNCHES = 64;

}  // namespace

struct hsh_handle {
    hsh_model m{};
    Plan plan;
    bool refused = false;
    bool host_only = false;
    int device = 0;
    hipStream_t stream = nullptr;
    // host copies in the reference layout
    std::vector<int32_t> ops;
    std::vector<double> chip;
    std::vector<double> tablist, xvals;
    // device buffers
    int32_t* d_prog = nullptr;
    BlockMeta* d_blocks = nullptr;
    TileMeta* d_tiles = nullptr;
    double *d_S = nullptr, *d_V = nullptr, *d_R = nullptr, *d_ta = nullptr, *d_tcf = nullptr,
           *d_tcb = nullptr, *d_tab = nullptr, *d_xvals = nullptr;
    int32_t *d_I = nullptr, *d_par = nullptr, *d_pred = nullptr, *d_spk = nullptr, *d_nspk = nullptr,
            *d_step = nullptr;
    DevModel dm{};
    JitModule jit;
    int batch_k = 32, batch_waves = 4;
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

void free_device(hsh_handle* h)
{
    void* ptrs[] = {h->d_prog, h->d_blocks, h->d_tiles, h->d_S, h->d_V, h->d_R, h->d_ta, h->d_tcf,
                    h->d_tcb, h->d_tab, h->d_xvals, h->d_I, h->d_par, h->d_pred, h->d_spk,
                    h->d_nspk, h->d_step};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    jit_free(h->jit);
    for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.second);
    for (auto& g : h->graph_src) (void)hipGraphDestroy(g.second);
    h->graphs.clear();
    h->graph_src.clear();
    for (hipEvent_t e : h->evs) (void)hipEventDestroy(e);
    h->evs.clear();
    if (h->ev_a) (void)hipEventDestroy(h->ev_a);
    if (h->ev_b) (void)hipEventDestroy(h->ev_b);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    h->stream = nullptr;
}

hsh_status batched_launch(hsh_handle* h, int K, int count)
{
    const int rows = h->plan.max_rows;
    if (rows == 1) {
        HIPCHK(launch_iso(h->dm, &h->jit, K, h->stream));
    } else {
        HIPCHK(launch_batched(h->dm, &h->jit, K, h->batch_waves, rows, h->stream));
    }
    (void)count;
    return HSH_OK;
}

hsh_status run_batched(hsh_handle* h, int nsteps, int* launches)
{
    const int K = std::max(1, h->batch_k);
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
            for (int i = 0; i < M && st == HSH_OK; i++) st = batched_launch(h, K, 1);
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
        hsh_status st = batched_launch(h, rem, 1);
        if (st != HSH_OK) return st;
        (*launches)++;
    }
    return HSH_OK;
}

}  // namespace

extern "C" {

const char* hsh_last_error(void) { return g_err.c_str(); }

hsh_status hsh_create(const hsh_model* m, const hsh_options* opt, hsh_handle** out)
{
    if (!m || !out) return fail(HSH_ERR_ARG, "hsh_create: null argument");
    *out = nullptr;
    if (m->ncompts <= 0 || m->nops <= 0 || m->nchips <= 0 || !m->ops || !m->chip || !m->vm ||
        !m->funcs || m->nfuncs <= 0)
        return fail(HSH_ERR_ARG, "hsh_create: empty model or missing ops/chip/vm/funcs");
    if (m->ncols > 0 && (!m->tablist || !m->xvals || m->xdivs <= 0))
        return fail(HSH_ERR_ARG, "hsh_create: ncols > 0 but no table");
    if (m->chanmode != 4 && m->chanmode != 5)
        return fail(HSH_ERR_ARG, "hsh_create: chanmode must be 4 or 5");
    if (m->method != HSH_BEULER && m->method != HSH_CRANK)
        return fail(HSH_ERR_ARG, "hsh_create: method must be HSH_BEULER or HSH_CRANK");
    hsh_handle* h = new hsh_handle();
    h->m = *m;
    h->ops.assign(m->ops, m->ops + m->nops);
    h->chip.assign(m->chip, m->chip + m->nchips);
    std::string err;
    if (!analyse(*m, h->plan, err)) {
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
    if (opt && (opt->flags & HSH_F_HOST_ONLY)) {
        h->host_only = true;
        if (opt->batch_k > 0) h->batch_k = opt->batch_k;
        *out = h;
        return HSH_OK;
    }
    h->device = opt ? opt->device : 0;
    if (opt && opt->batch_k > 0) h->batch_k = opt->batch_k;
    // 0 = four-wave tiles with register-held tree coefficients (trees of up
    // to 16 rows); otherwise waves per tile of the LDS-coefficient kernel
    if (opt && opt->batch_waves > 0) h->batch_waves = opt->batch_waves;
    else h->batch_waves = h->plan.max_rows <= 16 ? 0 : std::min(16, std::max(1, h->plan.max_rows / 4));
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
    HIPCHK_C(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    const Plan& P = h->plan;
    const size_t npos = P.blocks.size() * W;
    HIPCHK_C(dalloc(&h->d_prog, P.prog.size()));
    HIPCHK_C(dalloc(&h->d_blocks, P.blocks.size()));
    HIPCHK_C(dalloc(&h->d_tiles, P.tiles.size()));
    HIPCHK_C(dalloc(&h->d_S, (size_t)P.nS));
    HIPCHK_C(dalloc(&h->d_I, (size_t)std::max<int64_t>(1, P.nI)));
    HIPCHK_C(dalloc(&h->d_V, npos));
    HIPCHK_C(dalloc(&h->d_R, 2 * npos));
    HIPCHK_C(dalloc(&h->d_par, npos));
    HIPCHK_C(dalloc(&h->d_ta, npos));
    HIPCHK_C(dalloc(&h->d_tcf, npos));
    HIPCHK_C(dalloc(&h->d_tcb, npos));
    HIPCHK_C(dalloc(&h->d_pred, npos));
    HIPCHK_C(dalloc(&h->d_spk, 2 * (size_t)SPIKE_CAP));
    HIPCHK_C(dalloc(&h->d_nspk, 1));
    HIPCHK_C(dalloc(&h->d_step, P.tiles.size()));
    const size_t ntab = (m->ncols > 0) ? (size_t)(m->xdivs + 2) * m->ncols : 1;
    const size_t nx = (m->ncols > 0) ? (size_t)(m->xdivs + 2) : 1;
    HIPCHK_C(dalloc(&h->d_tab, ntab));
    HIPCHK_C(dalloc(&h->d_xvals, nx));
    // pack tree and predecessor arrays
    std::vector<int32_t> par(npos), pred(npos);
    std::vector<double> ta(npos), tcf(npos), tcb(npos);
    {
        std::vector<int32_t> lo_of(m->ncompts);
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
    HIPCHK_C(hipMemcpy(h->d_prog, P.prog.data(), P.prog.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK_C(hipMemcpy(h->d_blocks, P.blocks.data(), P.blocks.size() * sizeof(BlockMeta), hipMemcpyHostToDevice));
    HIPCHK_C(hipMemcpy(h->d_tiles, P.tiles.data(), P.tiles.size() * sizeof(TileMeta), hipMemcpyHostToDevice));
    HIPCHK_C(hipMemcpy(h->d_par, par.data(), npos * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK_C(hipMemcpy(h->d_pred, pred.data(), npos * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK_C(hipMemcpy(h->d_ta, ta.data(), npos * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK_C(hipMemcpy(h->d_tcf, tcf.data(), npos * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK_C(hipMemcpy(h->d_tcb, tcb.data(), npos * sizeof(double), hipMemcpyHostToDevice));
    if (m->ncols > 0) {
        HIPCHK_C(hipMemcpy(h->d_tab, m->tablist, ntab * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK_C(hipMemcpy(h->d_xvals, m->xvals, nx * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK_C(hipMemset(h->d_nspk, 0, sizeof(int32_t)));
    HIPCHK_C(hipMemset(h->d_step, 0, P.tiles.size() * sizeof(int32_t)));
    HIPCHK_C(hipMemset(h->d_R, 0, 2 * npos * sizeof(double)));
    if (upload_state(h, m->vm) != HSH_OK) return cleanup_fail(HSH_ERR_DEVICE);
    // reverse map chip index -> S positions
    {
        std::vector<int64_t> cnt(m->nchips + 1, 0);
        for (int g = 0; g < m->ncompts; g++) {
            const Template& t = P.templates[P.comp_tmpl[g]];
            for (int k = 0; k < t.nslots; k++) {
                int64_t src = P.dsrc[P.comp_dsrc_off[g] + k];
                if (src >= 0) cnt[src + 1]++;
            }
        }
        for (int i = 0; i < m->nchips; i++) cnt[i + 1] += cnt[i];
        h->c2s_off = cnt;
        h->c2s.assign(cnt[m->nchips], 0);
        std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
        for (int g = 0; g < m->ncompts; g++) {
            const Template& t = P.templates[P.comp_tmpl[g]];
            const int pos = P.comp_pos[g];
            const BlockMeta& bm = P.blocks[pos / W];
            for (int k = 0; k < t.nslots; k++) {
                int64_t src = P.dsrc[P.comp_dsrc_off[g] + k];
                if (src >= 0) h->c2s[cur[src]++] = (int64_t)bm.dslot + (int64_t)k * W + pos % W;
            }
        }
    }
    DevModel& d = h->dm;
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
    d.spk_cap = SPIKE_CAP;
    d.ncols = m->ncols;
    d.xdivs = m->xdivs;
    d.crank = (m->method == HSH_CRANK);
    d.xmin = m->xmin;
    d.invdx = m->invdx;
    d.nblocks = (int32_t)P.blocks.size();
    d.ntiles = (int32_t)P.tiles.size();
    d.last_pos = P.comp_pos[m->ncompts - 1];
    d.single = (m->ncompts == 1);
    d.max_slots = 0;
    for (const Template& t : P.templates) d.max_slots = std::max(d.max_slots, t.nslots);
    HIPCHK_C(hipEventCreate(&h->ev_a));
    HIPCHK_C(hipEventCreate(&h->ev_b));
#undef HIPCHK_C
    // kernels specialised to this model's channel programs; on failure the
    // built-in interpreter kernels run (message kept in hsh_last_error)
    const char* nojit = std::getenv("HSH_NO_JIT");
    if (!(opt && (opt->flags & HSH_F_NO_JIT)) && !(nojit && nojit[0] == '1')) {
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

hsh_status hsh_refusals(hsh_handle* h, hsh_refusal* out, int32_t cap, int32_t* n)
{
    if (!h || !n) return fail(HSH_ERR_ARG, "hsh_refusals: null argument");
    const auto& r = h->plan.refusals;
    *n = (int32_t)r.size();
    if (out)
        for (int32_t i = 0; i < cap && i < (int32_t)r.size(); i++) out[i] = r[i];
    return HSH_OK;
}

hsh_status hsh_step(hsh_handle* h, int32_t nsteps, int32_t mode, hsh_step_out* out)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_step: null handle");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_step: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_step: handle refused the model; use the CPU path");
    if (nsteps < 0) return fail(HSH_ERR_ARG, "hsh_step: negative step count");
    hsh_step_out o{};
    if (h->timing) HIPCHK(hipEventRecord(h->ev_a, h->stream));
    h->nev_used = 0;
    if (mode == HSH_PLAIN || mode == HSH_CHANNELS ||
        (mode == HSH_BATCHED && h->plan.max_rows > 64)) {
        for (int s = 0; s < nsteps; s++) {
            if (h->timing) {
                while ((int)h->evs.size() < h->nev_used + 3) {
                    hipEvent_t e;
                    HIPCHK(hipEventCreate(&e));
                    h->evs.push_back(e);
                }
                HIPCHK(hipEventRecord(h->evs[h->nev_used], h->stream));
            }
            HIPCHK(launch_channel(h->dm, &h->jit, h->stream));
            if (h->timing) HIPCHK(hipEventRecord(h->evs[h->nev_used + 1], h->stream));
            if (mode == HSH_CHANNELS) {
                HIPCHK(launch_step_advance(h->dm, 1, h->stream));
            } else {
                HIPCHK(launch_elim(h->dm, h->plan.max_rows, h->stream));
            }
            if (h->timing) {
                HIPCHK(hipEventRecord(h->evs[h->nev_used + 2], h->stream));
                h->nev_used += 3;
            }
            o.launches += 2;
        }
        o.vm_ready = (mode != HSH_CHANNELS);
    } else if (mode == HSH_BATCHED) {
        hsh_status st = run_batched(h, nsteps, &o.launches);
        if (st != HSH_OK) return st;
        o.vm_ready = 1;
    } else {
        return fail(HSH_ERR_ARG, "hsh_step: unknown mode");
    }
    if (h->timing) HIPCHK(hipEventRecord(h->ev_b, h->stream));
    h->total_valid = h->timing;
    h->host_step += nsteps;
    if (out) *out = o;
    return HSH_OK;
}

hsh_status hsh_sync(hsh_handle* h)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_sync: null handle");
    if (h->refused || h->host_only) return HSH_OK;
    HIPCHK(hipStreamSynchronize(h->stream));
    return HSH_OK;
}

hsh_status hsh_get_state(hsh_handle* h, double* vm, double* chip, double* results, int32_t* ops)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_get_state: null handle");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_get_state: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_get_state: handle refused the model");
    const Plan& P = h->plan;
    const size_t npos = P.blocks.size() * W;
    HIPCHK(hipStreamSynchronize(h->stream));
    if (vm || results) {
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
    }
    if (chip || ops) {
        hsh_status st = download_chip_ops(h);
        if (st != HSH_OK) return st;
        if (chip) std::memcpy(chip, h->chip.data(), h->chip.size() * sizeof(double));
        if (ops) std::memcpy(ops, h->ops.data(), h->ops.size() * sizeof(int32_t));
    }
    return HSH_OK;
}

hsh_status hsh_put_state(hsh_handle* h, const double* vm, const double* chip, const int32_t* ops)
{
    if (!h) return fail(HSH_ERR_ARG, "hsh_put_state: null handle");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_put_state: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_put_state: handle refused the model");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (chip || ops) {
        // refresh the host copies with the device's mutable state first, so
        // passing only one of chip/ops keeps the other current
        hsh_status st = download_chip_ops(h);
        if (st != HSH_OK) return st;
    }
    if (chip) std::memcpy(h->chip.data(), chip, h->chip.size() * sizeof(double));
    if (ops) std::memcpy(h->ops.data(), ops, h->ops.size() * sizeof(int32_t));
    if (chip || ops) {
        hsh_status st = upload_state(h, vm);
        return st;
    }
    if (vm) {
        std::vector<double> V;
        pack_V(h, vm, V);
        HIPCHK(hipMemcpy(h->d_V, V.data(), V.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return HSH_OK;
}

hsh_status hsh_set_chip(hsh_handle* h, int32_t n, const int32_t* idx, const double* val)
{
    if (!h || (n > 0 && (!idx || !val))) return fail(HSH_ERR_ARG, "hsh_set_chip: null argument");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_set_chip: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_set_chip: handle refused the model");
    std::vector<std::pair<int64_t, double>> w;
    for (int32_t i = 0; i < n; i++) {
        int32_t c = idx[i];
        if (c < 0 || c >= h->m.nchips) return fail(HSH_ERR_ARG, "hsh_set_chip: chip index out of range");
        h->chip[c] = val[i];
        for (int64_t q = h->c2s_off[c]; q < h->c2s_off[c + 1]; q++) w.emplace_back(h->c2s[q], val[i]);
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    for (auto& e : w)
        HIPCHK(hipMemcpy(h->d_S + e.first, &e.second, sizeof(double), hipMemcpyHostToDevice));
    return HSH_OK;
}

hsh_status hsh_drain_spikes(hsh_handle* h, int32_t* opidx, int32_t* step, int32_t cap, int32_t* n)
{
    if (!h || !n) return fail(HSH_ERR_ARG, "hsh_drain_spikes: null argument");
    if (h->host_only) return fail(HSH_ERR_STATE, "hsh_drain_spikes: host-only handle");
    if (h->refused) return fail(HSH_ERR_STATE, "hsh_drain_spikes: handle refused the model");
    HIPCHK(hipStreamSynchronize(h->stream));
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
    o.batch_k = h->batch_k;
    o.batch_waves = h->batch_waves;
    o.refused = h->refused ? 1 : 0;
    o.bytes_per_step = P.bytes_per_step;
    o.lane_fill = P.blocks.empty() ? 0.0 : (double)h->m.ncompts / ((double)P.blocks.size() * W);
    o.jit = h->jit.ok ? 1 : 0;
    o.jit_compile_s = h->jit.compile_s;
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

void hsh_destroy(hsh_handle* h)
{
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_device(h);
    delete h;
}

}  // extern "C"
