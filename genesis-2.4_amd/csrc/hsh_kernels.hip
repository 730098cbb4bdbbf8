// hsh_kernels.hip -- built-in gfx950 kernels of libhsolve_hip.
//
// Data layout (HBM): every array is [block][slot][lane] with 64 lanes per
// block, a block being one row of a tile of 64 neurons that share size and
// per-row channel programs (lane = neuron), or 64 same-program compartments
// of a pack of larger trees (lane = compartment).  A wave owns a block
// (channel programs) or a tile (elimination), so all loads of one slot by one
// wave are one contiguous 512-byte line and the program a wave interprets is
// wave-uniform.
//
// The channel kernels here run the generic micro-op interpreter; when the
// run-time specialised module of hsh_jit.cpp is available the launchers use
// its kernels instead (same bodies, straight-line programs).
#include "hsh_internal.h"
#include "hsh_device.h"

namespace hsh {

namespace {

template <bool REG, bool EXT>
__global__ __launch_bounds__(256) void k_channel(DevModel d)
{
    body_channel<InterpRun<REG, EXT>>(d);
}

// The basic variant is held to 128 VGPRs (four waves per SIMD): the C2
// population's 782 tiles then fit the chip in one round of workgroups.
template <int RPW, bool EXT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EXT ? 1 : 4))) void k_batched_r(DevModel d, int K)
{
    extern __shared__ double lds[];
    body_batched_r<RPW, InterpRun<true, EXT>>(d, K, lds);
}

template <bool REG, bool EXT>
__global__ __launch_bounds__(1024) void k_batched(DevModel d, int K)
{
    extern __shared__ double lds[];
    body_batched<InterpRun<REG, EXT>>(d, K, lds);
}

template <bool REG, bool EXT>
__global__ __launch_bounds__(256) void k_iso(DevModel d, int K)
{
    body_iso<InterpRun<REG, EXT>>(d, K);
}

template <bool REG, int STAGE, bool EXT>
__global__ __launch_bounds__(512) void k_forest(DevModel d, int K)
{
    extern __shared__ double lds[];
    body_forest<InterpRun<REG, EXT>, true, STAGE>(d, K, lds);
}

__global__ __launch_bounds__(256) void k_forest_elim(DevModel d)
{
    extern __shared__ double lds[];
    body_forest<InterpRun<false, false>, false, 0>(d, 1, lds);
}

// Plain drop-in, part (b): one lane per neuron, one wave per tile.  Forward
// elimination in push form -- when row r is final it is folded into its
// parent -- which applies each parent's kids in ascending row order, the
// order of the reference's funcs[] program; then back-substitution.
__global__ __launch_bounds__(256) void k_elim(DevModel d)
{
    int t = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= d.ntiles) return;
    const int lane = threadIdx.x & (W - 1);
    if (lane == 0) d.step[t] += 1;
    const TileMeta tm = d.tiles[t];
    const int R = tm.rows, b0 = tm.block0;
    double* Rr = d.R + (size_t)b0 * 2 * W + lane;     // row r: rhs Rr[r*2W], diag Rr[r*2W+W]
    const size_t p0 = (size_t)b0 * W + lane;
    // The reference never stores the eliminated diagonal of its last row
    // (hines_solve.c:275-276 writes only the solution), so that slot keeps
    // the channel update's value.
    const bool is_last = (p0 + (size_t)(R - 1) * W) == (size_t)d.last_pos;
    const double droot0 = Rr[(size_t)(R - 1) * 2 * W + W];
    for (int r = 0; r < R - 1; r++) {
        const int p = d.par[p0 + (size_t)r * W];
        if (p >= 0) {
            const double dk = Rr[(size_t)r * 2 * W + W];
            const double bk = Rr[(size_t)r * 2 * W];
            const double tt = d.ta[p0 + (size_t)r * W] / dk;
            Rr[(size_t)p * 2 * W + W] -= d.tcf[p0 + (size_t)r * W] * tt;
            Rr[(size_t)p * 2 * W] -= bk * tt;
        }
    }
    for (int r = R - 1; r >= 0; r--) {
        const size_t q = p0 + (size_t)r * W;
        double bv = Rr[(size_t)r * 2 * W];
        const int p = d.par[q];
        if (p >= 0) bv -= d.tcb[q] * Rr[(size_t)p * 2 * W];
        const double x = bv / Rr[(size_t)r * 2 * W + W];
        // a one-row solver keeps rhs in results[0] (hines_solve.c:236-238)
        if (!d.single) Rr[(size_t)r * 2 * W] = x;
        if (d.crank) {
            const double v = d.V[q];
            d.V[q] = x + x - v;
        } else {
            d.V[q] = x;
        }
    }
    if (is_last) Rr[(size_t)(R - 1) * 2 * W + W] = droot0;
}

// Plain drop-in elimination for trees of up to 16 rows with the batched
// kernel's row owners: four waves per tile, wave w the rows 4w..4w+3, their
// coefficients in registers, rhs/diag/V of the tile in LDS.
struct NoPrograms {
    static __device__ __forceinline__ CompOut run(const DevModel&, const BlockMeta&, int, double, bool, double,
                                                  bool, int)
    {
        return CompOut{0.0, 0.0};
    }
};

__global__ __launch_bounds__(256) void k_elim_w(DevModel d)
{
    extern __shared__ double lds[];
    body_batched_r<4, NoPrograms, false>(d, 1, lds);
}

// Plain drop-in elimination for trees of 17..ELIM_LDS_ROWS rows: one wave
// per tile with the tile's rhs/diag in LDS (M = rows of the largest tile sets
// the stride), so the parent updates of the sweep wait on LDS rather than on
// global memory.  The tree coefficients do not depend on the sweep: they are
// loaded ECH rows at a time into registers ahead of the rows that use them,
// so one global-memory latency is paid per chunk, not per row.  Same
// operations in the same order as k_elim.
constexpr int ECH = 16;

__global__ __launch_bounds__(64) void k_elim_l(DevModel d, int M)
{
    extern __shared__ double lds[];
    double* sB = lds;
    double* sD = lds + (size_t)M * W;
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    const TileMeta tm = d.tiles[t];
    const int R = tm.rows, b0 = tm.block0;
    if (lane == 0) d.step[t] += 1;
    double* Rr = d.R + (size_t)b0 * 2 * W + lane;
    const size_t p0 = (size_t)b0 * W + lane;
    // 16 rows' loads in flight before their LDS stores wait for them
#pragma unroll 16
    for (int r = 0; r < R; r++) {
        sB[r * W + lane] = Rr[(size_t)r * 2 * W];
        sD[r * W + lane] = Rr[(size_t)r * 2 * W + W];
    }
    const bool is_last = (p0 + (size_t)(R - 1) * W) == (size_t)d.last_pos;
    const double droot0 = sD[(R - 1) * W + lane];
    for (int r0 = 0; r0 < R - 1; r0 += ECH) {
        int pp[ECH];
        double ca[ECH], cf[ECH];
#pragma unroll
        for (int i = 0; i < ECH; i++) {
            const int r = r0 + i;
            pp[i] = -1;
            ca[i] = cf[i] = 0.0;
            if (r < R - 1) {
                const size_t q = p0 + (size_t)r * W;
                pp[i] = d.par[q];
                ca[i] = d.ta[q];
                cf[i] = d.tcf[q];
            }
        }
#pragma unroll
        for (int i = 0; i < ECH; i++) {
            const int r = r0 + i;
            const int p = pp[i];
            if (p >= 0) {
                // all four LDS reads before the writes (sB and sD share one
                // allocation, so the compiler cannot reorder them itself)
                const double dr = sD[r * W + lane], br = sB[r * W + lane];
                const double dp = sD[p * W + lane], bp = sB[p * W + lane];
                const double tt = ca[i] / dr;
                sD[p * W + lane] = dp - cf[i] * tt;
                sB[p * W + lane] = bp - br * tt;
            }
        }
    }
    for (int r1 = R - 1; r1 >= 0; r1 -= ECH) {
        int pp[ECH];
        double cb[ECH], vo[ECH];
#pragma unroll
        for (int i = 0; i < ECH; i++) {
            const int r = r1 - i;
            pp[i] = -1;
            cb[i] = vo[i] = 0.0;
            if (r >= 0) {
                const size_t q = p0 + (size_t)r * W;
                pp[i] = d.par[q];
                cb[i] = d.tcb[q];
                if (d.crank) vo[i] = d.V[q];
            }
        }
#pragma unroll
        for (int i = 0; i < ECH; i++) {
            const int r = r1 - i;
            if (r >= 0) {
                const size_t q = p0 + (size_t)r * W;
                const int p = pp[i];
                const double dr = sD[r * W + lane];
                const double xp = p >= 0 ? sB[p * W + lane] : 0.0;
                double bv = sB[r * W + lane];
                if (p >= 0) bv -= cb[i] * xp;
                const double x = bv / dr;
                if (!d.single) sB[r * W + lane] = x;
                d.V[q] = d.crank ? x + x - vo[i] : x;
            }
        }
    }
    if (is_last) sD[(R - 1) * W + lane] = droot0;
#pragma unroll 8
    for (int r = 0; r < R; r++) {
        Rr[(size_t)r * 2 * W] = sB[r * W + lane];
        Rr[(size_t)r * 2 * W + W] = sD[r * W + lane];
    }
}

// The elimination of a solver of symmetric compartments: one thread per
// unit interprets its funcs[] as do_euler_hsolve / do_crank_hsolve do
// (hines_solve.c:140-296), with results[2r] and results[2r+1] at rhs and
// diag of row r's block, and a mutable copy of its ravals: COPY_ARRAY fills
// working slots from the ravals stream, FASTSIBARRAY_ELIM and SIBARRAY_ELIM
// update them as pivot rows are eliminated, FORWARD_ELIM / BACKWARD_ELIM
// read the stream (working slots included) in order.  The host checked every
// index and the row cursor's path (sibling_unit, hsh_ingest.cpp).
__global__ __launch_bounds__(64) void k_funcs(DevModel d, const FuncsUnitDev* __restrict__ units, int nu,
                                              const int32_t* __restrict__ ffuncs, double* __restrict__ fwork)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const FuncsUnitDev U = units[u];
    const int b0 = d.tiles[U.tile].block0;
    const int lane = U.lane;
    const int32_t* f = ffuncs + U.f0;
    double* raval = fwork + U.q0;
    if (lane == 0) d.step[U.tile] += 1;
    auto RB = [&](int r) -> double& { return d.R[(long long)(b0 + r) * 2 * W + lane]; };
    auto RD = [&](int r) -> double& { return d.R[(long long)(b0 + r) * 2 * W + W + lane]; };
    auto VM = [&](int r) -> double& { return d.V[(long long)(b0 + r) * W + lane]; };
    int fi = 0, ra = 0, row = 1;
    double temp = 0.0, rv = RB(1), dv = RD(1);
    for (;;) {
        const int op = f[fi++];
        if (op == 0) {                    // FORWARD_ELIM
            const int k = f[fi++] >> 1;
            temp = raval[ra++] / RD(k);
            dv -= raval[ra++] * temp;
            rv -= RB(k) * temp;
        } else if (op == 1 || op == 2) {  // SET_DIAG, SKIP_DIAG
            RB(row) = rv;
            RD(row) = dv;
            row += op;
            rv = RB(row);
            dv = RD(row);
        } else if (op == 3) {             // FASTSIBARRAY_ELIM
            raval[f[fi++]] -= raval[ra++] * temp;
        } else if (op == 4) {             // COPY_ARRAY
            raval[f[fi++]] = raval[ra++];
        } else if (op == 5) {             // SIBARRAY_ELIM
            raval[f[fi + 1]] -= raval[f[fi]] * temp;
            fi += 2;
        } else {                          // FINISH
            break;
        }
    }
    // the last row, then the back-substitution up the rows
    int vi = U.nn - 1;
    if (d.crank) {
        rv = rv / dv;
        VM(vi) = rv + rv - VM(vi);
    } else {
        rv = rv / dv;
        VM(vi) = rv;
    }
    RB(row) = rv;
    row--;
    rv = RB(row);
    for (;;) {
        const int op = f[fi++];
        if (op == 0) {                    // BACKWARD_ELIM
            rv -= raval[ra++] * RB(f[fi++] >> 1);
        } else if (op == 6) {             // CALC_RESULTS
            rv = rv / RD(row);
            vi--;
            VM(vi) = d.crank ? rv + rv - VM(vi) : rv;
            RB(row) = rv;
            row--;
            if (row >= 0) rv = RB(row);
        } else if (op == 5) {             // SIBARRAY_ELIM
            rv -= raval[f[fi]] * RB(f[fi + 1] >> 1);
            fi += 2;
        } else {                          // FINISH
            break;
        }
    }
}

__global__ void k_step_advance(int32_t* step, int32_t n, int32_t by)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) step[t] += by;
}

constexpr int BATCH_MAX_ROWS = 64;
constexpr int ELIM_LDS_ROWS = 128;   // k_elim_l: 2 x 128 x 64 doubles = 128 KiB of LDS

hipError_t module_launch(hipFunction_t f, unsigned gx, unsigned bx, unsigned lds, hipStream_t s,
                         const DevModel& d, const int* K)
{
    DevModel dd = d;
    int kk = K ? *K : 0;
    void* args1[] = {&dd};
    void* args2[] = {&dd, &kk};
    return hipModuleLaunchKernel(f, gx, 1, 1, bx, 1, 1, lds, s, K ? args2 : args1, nullptr);
}

}  // namespace

int batched_lds_bytes(int max_rows) { return 3 * max_rows * W * (int)sizeof(double); }

// Interpreter variant of a launch: REG when every program's slots fit the
// register copy; the basic op set (no EXT) only with REG, and only when no
// program needs more.
enum class Variant { RegBasic, RegExt, Mem };

static Variant variant_of(const DevModel& d)
{
    if (d.max_slots > MAXS) return Variant::Mem;
    return d.ext ? Variant::RegExt : Variant::RegBasic;
}

hipError_t launch_channel(const DevModel& d, const JitModule* jm, hipStream_t s)
{
    const unsigned grid = (unsigned)((d.nblocks + 3) / 4);
    if (jm && jm->ok) return module_launch(jm->channel, grid, 256, 0, s, d, nullptr);
    switch (variant_of(d)) {
    case Variant::RegBasic: hipLaunchKernelGGL((k_channel<true, false>), dim3(grid), dim3(256), 0, s, d); break;
    case Variant::RegExt: hipLaunchKernelGGL((k_channel<true, true>), dim3(grid), dim3(256), 0, s, d); break;
    case Variant::Mem: hipLaunchKernelGGL((k_channel<false, true>), dim3(grid), dim3(256), 0, s, d); break;
    }
    return hipGetLastError();
}

hipError_t launch_elim(const DevModel& d, int max_rows, hipStream_t s)
{
    if (max_rows <= 16) {
        hipLaunchKernelGGL(k_elim_w, dim3(d.ntiles), dim3(256), (unsigned)batched_lds_bytes(max_rows), s, d);
    } else if (max_rows <= ELIM_LDS_ROWS) {
        const unsigned lds = (unsigned)(2 * max_rows * W * sizeof(double));
        hipLaunchKernelGGL(k_elim_l, dim3(d.ntiles), dim3(64), lds, s, d, max_rows);
    } else {
        dim3 grid((d.ntiles + 3) / 4);
        hipLaunchKernelGGL(k_elim, grid, dim3(256), 0, s, d);
    }
    return hipGetLastError();
}

template <int RPW>
void launch_batched_r(const DevModel& d, int ksteps, unsigned lds, hipStream_t s)
{
    if (d.ext)
        hipLaunchKernelGGL((k_batched_r<RPW, true>), dim3(d.ntiles), dim3(256), lds, s, d, ksteps);
    else
        hipLaunchKernelGGL((k_batched_r<RPW, false>), dim3(d.ntiles), dim3(256), lds, s, d, ksteps);
}

hipError_t launch_batched(const DevModel& d, const JitModule* jm, int ksteps, int waves, int max_rows,
                          hipStream_t s)
{
    if (max_rows > BATCH_MAX_ROWS) return hipErrorInvalidValue;
    const Variant v = variant_of(d);
    const unsigned lds = (unsigned)batched_lds_bytes(max_rows);
    const bool jit = jm && jm->ok;
    if (waves == 0 && max_rows <= 16 && (v != Variant::Mem || jit)) {
        const int which = max_rows <= 4 ? 0 : (max_rows <= 8 ? 1 : 2);
        if (jit) return module_launch(jm->batched_r[which], (unsigned)d.ntiles, 256, lds, s, d, &ksteps);
        if (which == 0)
            launch_batched_r<1>(d, ksteps, lds, s);
        else if (which == 1)
            launch_batched_r<2>(d, ksteps, lds, s);
        else
            launch_batched_r<4>(d, ksteps, lds, s);
        return hipGetLastError();
    }
    if (waves < 1) waves = 1;
    if (waves > 16) waves = 16;
    const dim3 grid(d.ntiles), block(waves * W);
    if (jit) return module_launch(jm->batched, (unsigned)d.ntiles, (unsigned)(waves * W), lds, s, d, &ksteps);
    switch (v) {
    case Variant::RegBasic: hipLaunchKernelGGL((k_batched<true, false>), grid, block, lds, s, d, ksteps); break;
    case Variant::RegExt: hipLaunchKernelGGL((k_batched<true, true>), grid, block, lds, s, d, ksteps); break;
    case Variant::Mem: hipLaunchKernelGGL((k_batched<false, true>), grid, block, lds, s, d, ksteps); break;
    }
    return hipGetLastError();
}

hipError_t launch_iso(const DevModel& d, const JitModule* jm, int ksteps, hipStream_t s)
{
    const unsigned grid = (unsigned)((d.ntiles + 3) / 4);
    if (jm && jm->ok) return module_launch(jm->iso, grid, 256, 0, s, d, &ksteps);
    switch (variant_of(d)) {
    case Variant::RegBasic: hipLaunchKernelGGL((k_iso<true, false>), dim3(grid), dim3(256), 0, s, d, ksteps); break;
    case Variant::RegExt: hipLaunchKernelGGL((k_iso<true, true>), dim3(grid), dim3(256), 0, s, d, ksteps); break;
    case Variant::Mem: hipLaunchKernelGGL((k_iso<false, true>), dim3(grid), dim3(256), 0, s, d, ksteps); break;
    }
    return hipGetLastError();
}

template <int STAGE>
void launch_forest_v(const DevModel& d, Variant v, int ksteps, unsigned bx, int lds, hipStream_t s)
{
    const dim3 grid(d.npacks), block(bx);
    switch (v) {
    case Variant::RegBasic: hipLaunchKernelGGL((k_forest<true, STAGE, false>), grid, block, lds, s, d, ksteps); break;
    case Variant::RegExt: hipLaunchKernelGGL((k_forest<true, STAGE, true>), grid, block, lds, s, d, ksteps); break;
    case Variant::Mem: hipLaunchKernelGGL((k_forest<false, STAGE, true>), grid, block, lds, s, d, ksteps); break;
    }
}

hipError_t launch_forest(const DevModel& d, const JitModule* jm, int ksteps, int waves, int stage, int lds,
                         hipStream_t s)
{
    if (d.npacks <= 0) return hipSuccess;
    if (stage < 0 || stage > 2) return hipErrorInvalidValue;
    waves = waves < 1 ? 1 : (waves > 8 ? 8 : waves);
    const unsigned bx = (unsigned)(waves * W);
    if (jm && jm->ok) return module_launch(jm->forest[stage], (unsigned)d.npacks, bx, (unsigned)lds, s, d, &ksteps);
    if (stage == 2)
        launch_forest_v<2>(d, variant_of(d), ksteps, bx, lds, s);
    else if (stage == 1)
        launch_forest_v<1>(d, variant_of(d), ksteps, bx, lds, s);
    else
        launch_forest_v<0>(d, variant_of(d), ksteps, bx, lds, s);
    return hipGetLastError();
}

hipError_t launch_forest_elim(const DevModel& d, int lds, hipStream_t s)
{
    if (d.npacks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_forest_elim, dim3(d.npacks), dim3(256), lds, s, d);
    return hipGetLastError();
}

hipError_t launch_funcs(const DevModel& d, const FuncsUnitDev* u, int nu, const int32_t* ffuncs, double* fwork,
                        hipStream_t s)
{
    if (nu <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_funcs, dim3((nu + 63) / 64), dim3(64), 0, s, d, u, nu, ffuncs, fwork);
    return hipGetLastError();
}

hipError_t launch_step_advance(const DevModel& d, int32_t by, hipStream_t s)
{
    const int n = d.ntiles + d.npacks;
    hipLaunchKernelGGL(k_step_advance, dim3((n + 255) / 256), dim3(256), 0, s, d.step, n, by);
    return hipGetLastError();
}

}  // namespace hsh
