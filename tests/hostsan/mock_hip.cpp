// Mock HIP runtime + kernel-launch stubs for an AddressSanitizer build of the
// host code.  Every launch checks the pointers of its DevModel against live
// mock allocations (with the extents the grid implies for V, R, blocks,
// tiles, packs, step counters); captured graphs keep their launches'
// DevModels and re-check them at every replay.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "hsh_internal.h"
#include <hip/hiprtc.h>

using namespace hsh;

static std::map<uintptr_t, size_t> g_live;
static size_t g_nstatic = 0;     // entries of g_live that are static device arrays
static std::mutex g_mu;
static long g_nlaunch = 0, g_nreplay = 0;

[[noreturn]] static void die(const char* what, const char* name, const void* p, size_t n)
{
    std::fprintf(stderr, "MOCKHIP ERROR: %s: %s ptr=%p bytes=%zu\n", what, name, p, n);
    std::fflush(stderr);
    std::abort();
}

// 0: [p, p+n) lies in one live allocation; 1: p is in none; 2: the range
// runs past the end of the allocation holding p
static int range_state(const void* p, size_t n)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto u = (uintptr_t)p;
    auto it = g_live.upper_bound(u);
    if (it == g_live.begin()) return 1;
    --it;
    const uintptr_t end = it->first + it->second;
    if (u > end || (n > 0 && u == end)) return 1;       // (an empty range may sit at the end)
    return u + n > end ? 2 : 0;
}

static void check_range(const char* name, const void* p, size_t n)
{
    if (!p) return;
    const int st = range_state(p, n);
    if (st == 1) die("pointer outside every live allocation", name, p, n);
    if (st == 2) die("range outside its allocation", name, p, n);
}

static void check_model(const DevModel& d, const char* who)
{
    const size_t nb = (size_t)std::max(0, d.nblocks);
    check_range("S", d.S, 8);
    check_range("V", d.V, nb * W * 8);
    check_range("R", d.R, nb * 2 * W * 8);
    check_range("blocks", d.blocks, nb * sizeof(BlockMeta));
    check_range("tiles", d.tiles, (size_t)std::max(0, d.ntiles) * sizeof(TileMeta));
    check_range("packs", d.packs, (size_t)std::max(0, d.npacks) * sizeof(PackMeta));
    check_range("step", d.step, (size_t)std::max(1, d.ntiles + d.npacks) * 4);
    check_range("prog", d.prog, 4);
    check_range("I", d.I, 4);
    check_range("par", d.par, nb * W * 4);
    check_range("ta", d.ta, nb * W * 8);
    check_range("tcf", d.tcf, nb * W * 8);
    check_range("tcb", d.tcb, nb * W * 8);
    check_range("pred", d.pred, 4);
    check_range("tab", d.tab, 8);
    check_range("xvals", d.xvals, 8);
    check_range("spk", d.spk, 8);
    check_range("nspk", d.nspk, 4);
    check_range("pstep", d.pstep, (size_t)std::max(0, d.npacks) * 4);
    check_range("lvl", d.lvl, 4);
    check_range("fent", d.fent, 4);
    check_range("kidx", d.kidx, 2);
    check_range("kta", d.kta, 8);
    check_range("ktcf", d.ktcf, 8);
    check_range("bkp", d.bkp, 4);
    check_range("btcb", d.btcb, 8);
    check_range("stab", d.stab, 8);
    check_range("giv", d.giv, 8);
    check_range("evctl", d.evctl, sizeof(EvCtl));
    (void)who;
}

// ---- streams, capture, graphs
struct MockStream {
    bool capturing = false;
    std::vector<DevModel> cap;
};
struct MockGraph {
    std::vector<DevModel> launches;
};
struct MockGraphExec {
    std::vector<DevModel> launches;
};
struct MockEvent {
    int id;
};

static void record(const DevModel& d, hipStream_t s, const char* who)
{
    check_model(d, who);
    g_nlaunch++;
    if (s && s->capturing) s->cap.push_back(d);
}

hipError_t hipMalloc(void** p, size_t n)
{
    void* q = std::malloc(n ? n : 1);
    if (!q) return hipErrorOutOfMemory;
    std::memset(q, 0xA5, n);
    std::lock_guard<std::mutex> lk(g_mu);
    g_live[(uintptr_t)q] = n ? n : 1;
    *p = q;
    return hipSuccess;
}
hipError_t hipFree(void* p)
{
    if (!p) return hipSuccess;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_live.find((uintptr_t)p);
        if (it == g_live.end()) die("hipFree of a pointer not allocated", "free", p, 0);
        g_live.erase(it);
    }
    std::free(p);
    return hipSuccess;
}
static hipError_t copy(void* d, const void* s, size_t n, hipMemcpyKind k)
{
    if (n == 0) return hipSuccess;
    if (k == hipMemcpyHostToDevice || k == hipMemcpyDeviceToDevice) check_range("memcpy dst", d, n);
    if (k == hipMemcpyDeviceToHost || k == hipMemcpyDeviceToDevice) check_range("memcpy src", s, n);
    std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k) { return copy(d, s, n, k); }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st)
{
    if (st && st->capturing) die("memcpy inside capture", "memcpy", d, n);
    return copy(d, s, n, k);
}
hipError_t hipMemset(void* d, int v, size_t n)
{
    check_range("memset", d, n);
    std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st)
{
    if (st && st->capturing) die("memset inside capture", "memset", d, n);
    return hipMemset(d, v, n);
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = new MockStream(); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t s)
{
    if (s && s->capturing) die("synchronize inside capture", "sync", s, 0);
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode)
{
    if (s->capturing) die("nested capture", "capture", s, 0);
    s->capturing = true;
    s->cap.clear();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g)
{
    if (!s->capturing) die("end of a capture not begun", "capture", s, 0);
    s->capturing = false;
    *g = new MockGraph{s->cap};
    s->cap.clear();
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t)
{
    *e = new MockGraphExec{g->launches};
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s)
{
    if (s && s->capturing) die("graph launch inside capture", "graph", e, 0);
    for (const DevModel& d : e->launches) check_model(d, "graph replay");
    g_nreplay++;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }

static int g_evid = 0;
static unsigned long long g_rng = 0;
hipError_t hipEventCreate(hipEvent_t* e) { *e = new MockEvent{g_evid++}; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t)
{
    // pseudo-random timings (MOCKHIP_SEED) so that different candidates win
    if (!g_rng) {
        const char* s = std::getenv("MOCKHIP_SEED");
        g_rng = 0x9E3779B97F4A7C15ull ^ (unsigned long long)(s ? std::atoll(s) : 1);
    }
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    *ms = 1.0f + (float)(g_rng % 1000) / 1000.0f;
    return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "mock error"; }
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int)
{
    std::memset(p, 0, sizeof *p);
    std::strcpy(p->name, "mock");
    std::strcpy(p->gcnArchName, "gfx950:sramecc+:xnack-");
    p->multiProcessorCount = 256;
    p->totalGlobalMem = (size_t)288 << 30;
    p->sharedMemPerBlock = 160 * 1024;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int)
{
    *v = a == hipDeviceAttributeMultiprocessorCount ? 256 : 160 * 1024;
    return hipSuccess;
}
struct MockModule {
    int functions = 0;
};
struct MockFunction {
    int id;
};
static long g_nmodules = 0;
static MockFunction g_functions[64];
hipError_t hipModuleLoadData(hipModule_t* m, const void* image)
{
    if (!image) return hipErrorInvalidValue;
    *m = new MockModule();
    g_nmodules++;
    return hipSuccess;
}
hipError_t hipModuleGetFunction(hipFunction_t* f, hipModule_t m, const char* name)
{
    if (!m || !name || !name[0]) return hipErrorInvalidValue;
    *f = &g_functions[m->functions++ % 64];
    return hipSuccess;
}
hipError_t hipModuleUnload(hipModule_t m)
{
    if (!m) return hipErrorInvalidValue;
    delete m;
    g_nmodules--;
    return hipSuccess;
}

// The run-time compiler: MOCKHIP_RTC_FAIL=1 fails every compilation (the
// library then runs its interpreter kernels); otherwise a program compiles
// to a small stand-in code object once its source is non-empty.
struct MockRtc {
    std::string src;
};
hiprtcResult hiprtcCreateProgram(hiprtcProgram* p, const char* src, const char*, int, const char**, const char**)
{
    *p = new MockRtc{src ? src : ""};
    return HIPRTC_SUCCESS;
}
hiprtcResult hiprtcCompileProgram(hiprtcProgram p, int, const char**)
{
    const char* f = std::getenv("MOCKHIP_RTC_FAIL");
    if ((f && f[0] == '1') || p->src.empty()) return HIPRTC_ERROR_COMPILATION;
    return HIPRTC_SUCCESS;
}
hiprtcResult hiprtcGetProgramLogSize(hiprtcProgram, size_t* n) { *n = 1; return HIPRTC_SUCCESS; }
hiprtcResult hiprtcGetProgramLog(hiprtcProgram, char* log) { log[0] = 0; return HIPRTC_SUCCESS; }
hiprtcResult hiprtcDestroyProgram(hiprtcProgram* p) { delete *p; *p = nullptr; return HIPRTC_SUCCESS; }
hiprtcResult hiprtcGetCodeSize(hiprtcProgram, size_t* n) { *n = 64; return HIPRTC_SUCCESS; }
hiprtcResult hiprtcGetCode(hiprtcProgram, char* code) { std::memset(code, 0, 64); return HIPRTC_SUCCESS; }
const char* hiprtcGetErrorString(hiprtcResult) { return "mock compilation failure"; }

extern "C" long mockhip_launches() { return g_nlaunch; }
extern "C" long mockhip_replays() { return g_nreplay; }
extern "C" long mockhip_live() { std::lock_guard<std::mutex> lk(g_mu); return (long)(g_live.size() - g_nstatic); }
extern "C" long mockhip_modules() { return g_nmodules; }
// The checks themselves: a range inside a live allocation passes, one that
// runs past its end or starts in a freed one does not.  0 = all as expected.
extern "C" int mockhip_selfcheck()
{
    void* p = nullptr;
    if (hipMalloc(&p, 64) != hipSuccess) return 1;
    int bad = 0;
    bad |= range_state(p, 64) != 0;
    bad |= range_state((char*)p + 32, 32) != 0;
    bad |= (range_state((char*)p + 32, 64) != 2) << 1;
    bad |= (range_state((char*)p + 64, 8) != 1) << 2;
    hipFree(p);
    bad |= (range_state(p, 8) != 1) << 3;
    return bad;
}

// ---- kernel launchers (hsh_kernels.hip, hsh_events.hip, hsh_adjoint.hip)
namespace hsh {
static InEntry g_static_inputs[16384];
int batched_lds_bytes(int max_rows) { return 3 * max_rows * W * (int)sizeof(double); }
hipError_t launch_channel(const DevModel& d, const JitModule*, hipStream_t s) { record(d, s, "channel"); return hipSuccess; }
hipError_t launch_elim(const DevModel& d, int, hipStream_t s) { record(d, s, "elim"); return hipSuccess; }
hipError_t launch_batched(const DevModel& d, const JitModule*, int, int, int, hipStream_t s)
{
    record(d, s, "batched");
    return hipSuccess;
}
hipError_t launch_iso(const DevModel& d, const JitModule*, int, hipStream_t s) { record(d, s, "iso"); return hipSuccess; }
hipError_t launch_forest(const DevModel& d, const JitModule*, int, int, int stage, int lds, hipStream_t s)
{
    if (lds > 160 * 1024) die("forest LDS over 160 KiB", "lds", nullptr, (size_t)lds);
    if (stage < 0 || stage > 2) die("forest staging mode out of range", "stage", nullptr, (size_t)stage);
    record(d, s, "forest");
    return hipSuccess;
}
hipError_t launch_forest_elim(const DevModel& d, int, hipStream_t s) { record(d, s, "forest_elim"); return hipSuccess; }
hipError_t launch_step_advance(const DevModel& d, int32_t, hipStream_t s) { record(d, s, "advance"); return hipSuccess; }
hipError_t launch_funcs(const DevModel& d, const FuncsUnitDev* u, int nu, const int32_t* f, double* w, hipStream_t s)
{
    check_range("funcs units", u, (size_t)nu * sizeof(FuncsUnitDev));
    check_range("ffuncs", f, 4);
    check_range("fwork", w, 8);
    record(d, s, "funcs");
    return hipSuccess;
}
hipError_t launch_ev_spikes(const DevModel& d, hipStream_t s) { record(d, s, "ev_spikes"); return hipSuccess; }
hipError_t launch_ev_post(const DevModel& d, const EvPost&, EvCtl*, hipStream_t s) { record(d, s, "ev_post"); return hipSuccess; }
hipError_t launch_inputs(double* S, const InEntry* e, int n, hipStream_t)
{
    check_range("inputs S", S, 8);
    if (e != g_static_inputs) check_range("inputs", e, (size_t)n * sizeof(InEntry));
    return hipSuccess;
}
// the library's static staging array: a device array of program lifetime
hipError_t inputs_static_setup(int)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_live.emplace((uintptr_t)g_static_inputs, sizeof g_static_inputs).second) g_nstatic++;
    return hipSuccess;
}
InEntry* inputs_static_ptr(int) { return g_static_inputs; }
hipError_t launch_adj_rhs(const AdjModel& a, hipStream_t) { check_model(a.d, "adj"); return hipSuccess; }
hipError_t launch_adj_elim(const AdjModel& a, hipStream_t) { check_model(a.d, "adj"); return hipSuccess; }
hipError_t launch_adj_local(const AdjModel& a, hipStream_t) { check_model(a.d, "adj"); return hipSuccess; }
hipError_t launch_adj_forest(const AdjPack& a, const JitModule*, int, hipStream_t) { check_model(a.d, "adjf"); return hipSuccess; }
hipError_t launch_gather(const double* V, const int* pos, int n, double* out, hipStream_t)
{
    check_range("gather V", V, 8);
    check_range("gather pos", pos, (size_t)n * 4);
    check_range("gather out", out, (size_t)n * 8);
    return hipSuccess;
}
hipError_t launch_inject(double* LV, const int* pos, int n, const double* g, hipStream_t)
{
    check_range("inject LV", LV, 8);
    check_range("inject pos", pos, (size_t)n * 4);
    check_range("inject g", g, (size_t)n * 8);
    return hipSuccess;
}
}  // namespace hsh
