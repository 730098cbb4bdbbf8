// hsh_events.hip -- the step's spike pass and the posting of outside events
// into the synchans' pending lists.
//
// The reference delivers a spike while it walks ops[]: SPIKE_OP calls
// h_dospike_event (hines_conn.c:44-56), each connected synchan's EVENT puts
// the spike into its pending list (newconn/synchan.c:161-302), and a SYN2_OP
// later in the same walk may already see the list changed.  A spike can hit
// as early as the next step: every EVENT re-times the list head by the whole
// steps elapsed since time_last_event, truncating the floating interval, so
// the head's time field drifts upward and events queued behind it fire
// early.  So the device decides all of a step's spikes first -- a SPIKE_OP
// reads only Vm (the last step's solution), its threshold and its counter --
// and every SYN2_OP then runs its synchan's EVENTs and fire in ops[] order
// (syn2_op, hsh_device.h).
#include "hsh_internal.h"
#include "hsh_device.h"

#include <mutex>

namespace hsh {
namespace {

// Staging of hsh_set_inputs_static: device memory of fixed size with static
// storage, so that entry point allocates nothing.  Each GPU holds its own
// copy until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (scratch_ptr).  Every
// call on a GPU shares that GPU's array; a call reads only the entries it
// has just written, so nothing in it needs clearing.
static __device__ __align__(256) InEntry scratch[HSH_INPUTS_STATIC_MAX];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
InEntry* scratch_ptr[MAX_DEVICES] = {};

// The step's spike pass for every spike source (spike_src_pass), and the
// clock of the step.  A spike adds each of its connections to its target
// synchan's arrivals, which the synchan takes in slot order when its SYN2_OP
// runs.  Every step counter holds the step about to run.
__global__ __launch_bounds__(256) void k_ev_spikes(DevModel d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) ev_clock(d, d.step[0]);
    const bool live = i < d.nsrc;
    SpikeSrc q{};
    double v = 0.0;
    int step = 0;
    if (live) {
        q = d.src[i];
        v = d.V[q.vpos];
        step = d.step[d.blocks[q.vpos / W].tile];
    }
    spike_src_pass(d, live, q, v, step);
}

// Events posted from the host (hsh_post_events), one thread per synchan:
// those timed after the last step's ops first, then those before the next
// step's, each an EVENT at that step's SimulationTime().
__global__ __launch_bounds__(256) void k_ev_post(DevModel d, EvPost p, EvCtl* ctl)
{
    const int sid = blockIdx.x * blockDim.x + threadIdx.x;
    if (sid >= p.nsyn) return;
    const int np = ctl->npost;
    int lo = 0, hi = np;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (p.sid[mid] < sid) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= np || p.sid[lo] != sid) return;
    const int off = d.lst_off[sid], cap = d.lst_cap[sid];
    int n = d.lst_n[sid];
    int cnt = d.I[p.ipos[sid]];
    double tle = d.tle[sid];
    bool full = false;
    for (int pass = 1; pass >= 0; pass--)
        for (int i = lo; i < np && p.sid[i] == sid; i++)
            if (p.after[i] == pass)
                full |= !ev_insert(d.lst_t + off, d.lst_w + off, n, cap, cnt, tle, p.T[i], p.wdt[i],
                                   pass ? ctl->tnow : ctl->tnext, d.dt);
    d.lst_n[sid] = n;
    d.tle[sid] = tle;
    d.I[p.ipos[sid]] = cnt;
    if (full) atomicOr(&ctl->err, EV_ERR_FULL);
}

// Message inputs: the first entry of each position applies that position's
// entries in order (a later set overwrites, adds accumulate in order).
__global__ __launch_bounds__(256) void k_inputs(double* __restrict__ S, const InEntry* __restrict__ e, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (i > 0 && e[i - 1].pos == e[i].pos)) return;
    const int pos = e[i].pos;
    double v = S[pos];
    for (int j = i; j < n && e[j].pos == pos; j++) v = e[j].mode ? v + e[j].value : e[j].value;
    S[pos] = v;
}

}  // namespace

hipError_t inputs_static_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_ptr[device]) return hipSuccess;
    // the caller has made `device` current
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    scratch_ptr[device] = static_cast<InEntry*>(p);
    return hipSuccess;
}

InEntry* inputs_static_ptr(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mu);
    return scratch_ptr[device];
}

hipError_t launch_inputs(double* S, const InEntry* e, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_inputs, dim3((n + 255) / 256), dim3(256), 0, s, S, e, n);
    return hipGetLastError();
}

hipError_t launch_ev_spikes(const DevModel& d, hipStream_t s)
{
    const int n = d.nsrc > 0 ? d.nsrc : 1;
    hipLaunchKernelGGL(k_ev_spikes, dim3((n + 255) / 256), dim3(256), 0, s, d);
    return hipGetLastError();
}

hipError_t launch_ev_post(const DevModel& d, const EvPost& p, EvCtl* ctl, hipStream_t s)
{
    if (p.nsyn <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ev_post, dim3((p.nsyn + 255) / 256), dim3(256), 0, s, d, p, ctl);
    return hipGetLastError();
}

}  // namespace hsh
