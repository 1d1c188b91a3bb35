// Persistent batched SIRT over several geometries in one launch, and hipGraph capture
// of SIRT iterations.
//
// A batch holds members, each one distance-driven plan (npix, angles) with its own stack
// of slices.  One iteration is two persistent launches (forward, back): a fixed grid of
// blocks pulls work items -- 64 bins of one view of one slice (one 64-thread block), or
// a 64x64 tile of one slice (256 threads) -- from an atomic counter, largest items
// first, and runs the same per-tile device code as the per-plan kernels.  Every output element is still
// written by one thread in the reference's accumulation order, so results do not depend
// on the schedule and equal the per-plan path bit for bit.
#include <algorithm>
#include <new>
#include <vector>

#include "dd_device.h"

namespace itomo {

struct BatchMember {
    const DdView* geom;
    const float* lstar;
    int nang, npix, nslices;
    float* x;
    const float* b;
    const float* R;
    const float* C;
    float* work;
    float* xT;
    double* pf;
    double* pb;
    int bps_f, bps_b;
};

struct BatchItem {
    int member;
    int a;  // forward: bin block; back: tile column
    int b;  // forward: view;      back: tile row
    int z;
};

struct SliceRef {   // for the finalize kernel
    const double* pf;
    const double* pb;
    int bps_f, bps_b;
    double* stats;  // 3 doubles, or null
};

}  // namespace itomo

struct itomo_batch {
    int nmembers = 0;
    int total_slices = 0;
    int nfwd = 0, nback = 0;
    int grid = 0;
    std::vector<itomo::BatchMember> h_members;
    std::vector<const itomo_plan*> plans;
    itomo::BatchMember* d_members = nullptr;
    itomo::BatchItem* d_fwd = nullptr;
    itomo::BatchItem* d_back = nullptr;
    itomo::SliceRef* d_slices = nullptr;
    int* d_counter = nullptr;
    void* d_scratch = nullptr;
};

struct itomo_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* scratch = nullptr;   // the graph's own scratch region (single-plan graphs)
};

namespace itomo {

__global__ void reset_counter_kernel(int* c) { c[0] = 0; c[1] = 0; }

// Bins per forward work item (one 64-thread block): the per-plan tuner's choice at large
// sizes; smaller items also balance the persistent grid better than 256-bin ones.
constexpr int kBatchFwdBlock = 64;

__device__ __forceinline__ int next_item(int* counter)
{
    __shared__ int s_item;
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = s_item;
    __syncthreads();  // all threads have read s_item before it is overwritten
    return it;
}

// Five waves per SIMD (96 registers, three of them spilled): the kernel holds every
// forward path and the fused update, and left to itself the compiler takes 129
// registers, room for three waves.  Measured on the 1024^2 x 900 stack: 69.0 ms per
// iteration at four waves, 63.9 at five, 70.7 at six (18 registers spilled).
__global__ void __launch_bounds__(kBatchFwdBlock) __attribute__((amdgpu_waves_per_eu(5)))
batch_fwd_kernel(const BatchMember* __restrict__ mem, const BatchItem* __restrict__ items,
                 int nitems, int* counter)
{
    for (;;) {
        const int it = next_item(counter);
        if (it >= nitems) break;
        const BatchItem w = items[it];
        const BatchMember m = mem[w.member];
        EpiArgs e{};
        e.b = m.b;
        e.w = m.R;
        e.partials = m.pf;
        e.blocks_per_slice = m.bps_f;
        dd_fwd_block<ITOMO_ARITH_FAST, EPI_SIRT_FWD, false, kBatchFwdBlock>(
            m.geom, m.lstar, m.nang, m.npix, m.x, m.xT, m.work, (int64_t)m.npix * m.npix,
            (int64_t)m.nang * m.npix, 0, e, w.a, w.b, w.z);
    }
}

__global__ void __launch_bounds__(256)
batch_back_kernel(const BatchMember* __restrict__ mem, const BatchItem* __restrict__ items,
                  int nitems, int* counter, float pc_beta)
{
    for (;;) {
        const int it = next_item(counter + 1);
        if (it >= nitems) break;
        const BatchItem w = items[it];
        const BatchMember m = mem[w.member];
        EpiArgs e{};
        e.w = m.C;
        e.pc_beta = pc_beta;
        e.partials = m.pb;
        e.blocks_per_slice = m.bps_b;
        const int nbx = (m.npix + kDdBackTile - 1) / kDdBackTile;
        dd_back_tile<ITOMO_ARITH_FAST, EPI_SIRT_BACK, false, kDdBackTile, 4>(
            m.geom, m.lstar, m.nang, m.npix, m.work, m.x, (int64_t)m.nang * m.npix,
            (int64_t)m.npix * m.npix, 0, e, w.a, w.b, w.z, w.b * nbx + w.a);
    }
}

// One block per slice of the whole batch: sum the forward and back partials.
__global__ void __launch_bounds__(256)
batch_finalize_kernel(const SliceRef* __restrict__ refs)
{
    __shared__ double red[3][256];
    const SliceRef r = refs[blockIdx.x];
    if (!r.stats) return;
    const int t = threadIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int b = t; b < r.bps_f; b += 256) s0 += r.pf[(int64_t)b * 3 + 0];
    for (int b = t; b < r.bps_b; b += 256) {
        s1 += r.pb[(int64_t)b * 3 + 1];
        s2 += r.pb[(int64_t)b * 3 + 2];
    }
    red[0][t] = s0;
    red[1][t] = s1;
    red[2][t] = s2;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
            red[0][t] += red[0][t + off];
            red[1][t] += red[1][t + off];
            red[2][t] += red[2][t + off];
        }
        __syncthreads();
    }
    if (t == 0) {
        r.stats[0] = red[0][0];
        r.stats[1] = red[1][0];
        r.stats[2] = red[2][0];
    }
}

__global__ void __launch_bounds__(256)
transpose_kernel_b(const float* __restrict__ in, float* __restrict__ out, int npix,
                   int64_t stride)
{
    transpose_tile(in, out, npix, stride, blockIdx.x, blockIdx.y, blockIdx.z);
}

}  // namespace itomo

using namespace itomo;

extern "C" {

int itomo_batch_destroy(itomo_batch* B);

int itomo_batch_create(itomo_batch** out, int nmembers, const itomo_plan* const* plans,
                       const int* nslices, float* const* x, const float* const* b,
                       const float* const* R, const float* const* C, float* const* work,
                       double* const* stats)
{
    if (!out || nmembers < 1 || !plans || !nslices || !x || !b || !R || !C || !work)
        return set_error(ITOMO_EINVAL, "batch_create: bad argument");
    *out = nullptr;
    itomo_batch* B = new (std::nothrow) itomo_batch();
    if (!B) return set_error(ITOMO_ENOMEM, "batch_create: out of host memory");
    B->nmembers = nmembers;
    std::vector<BatchItem> fwd, back;
    std::vector<SliceRef> refs;
    std::vector<size_t> cost_f, cost_b;
    size_t scratch_bytes = 0;
    std::vector<size_t> off_xT(nmembers), off_pf(nmembers), off_pb(nmembers);
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    for (int m = 0; m < nmembers; ++m) {
        const itomo_plan* p = plans[m];
        if (!p || p->projector != ITOMO_PROJ_DD || p->arith != ITOMO_ARITH_FAST) {
            delete B;
            return set_error(ITOMO_EINVAL, "batch_create: member %d is not a distance-driven "
                                           "plan with the reference's fast arithmetic", m);
        }
        if (nslices[m] < 1 || nslices[m] > 65535 || !x[m] || !b[m] || !R[m] || !C[m] || !work[m]) {
            delete B;
            return set_error(ITOMO_EINVAL, "batch_create: member %d has bad arrays", m);
        }
        BatchMember bm{};
        bm.geom = (const DdView*)p->d_geom;
        bm.lstar = p->d_lstar;
        bm.nang = p->nang;
        bm.npix = p->npix;
        bm.nslices = nslices[m];
        bm.x = x[m];
        bm.b = b[m];
        bm.R = R[m];
        bm.C = C[m];
        bm.work = work[m];
        // the batch launches its own shapes whatever the plan's choice
        bm.bps_f = (int)dd_fwd_slots(p->npix, p->nang);
        bm.bps_b = (int)dd_back_blocks(p->npix, kDdBackTile);
        B->h_members.push_back(bm);
        B->plans.push_back(p);
        off_xT[m] = scratch_bytes;
        scratch_bytes += up(sizeof(float) * (size_t)p->npix * p->npix * nslices[m]);
        off_pf[m] = scratch_bytes;
        scratch_bytes += up(sizeof(double) * 3 * (size_t)bm.bps_f * nslices[m]);
        off_pb[m] = scratch_bytes;
        scratch_bytes += up(sizeof(double) * 3 * (size_t)bm.bps_b * nslices[m]);
        const int nbf = (p->npix + kBatchFwdBlock - 1) / kBatchFwdBlock;
        const int nbt = (p->npix + kDdBackTile - 1) / kDdBackTile;
        for (int z = 0; z < nslices[m]; ++z) {
            for (int v = 0; v < p->nang; ++v)
                for (int a = 0; a < nbf; ++a) fwd.push_back({m, a, v, z});
            for (int by = 0; by < nbt; ++by)
                for (int bx = 0; bx < nbt; ++bx) back.push_back({m, bx, by, z});
            refs.push_back({nullptr, nullptr, bm.bps_f, bm.bps_b,
                            stats && stats[m] ? stats[m] + 3 * z : nullptr});
        }
        B->total_slices += nslices[m];
    }
    // largest items first: forward items cost ~ npix (lines), back tiles ~ nang (views)
    std::stable_sort(fwd.begin(), fwd.end(), [&](const BatchItem& u, const BatchItem& v) {
        return B->h_members[u.member].npix > B->h_members[v.member].npix;
    });
    std::stable_sort(back.begin(), back.end(), [&](const BatchItem& u, const BatchItem& v) {
        return B->h_members[u.member].nang > B->h_members[v.member].nang;
    });
    B->nfwd = (int)fwd.size();
    B->nback = (int)back.size();
    int rc = check_hip(hipMalloc(&B->d_scratch, scratch_bytes), "batch_create: hipMalloc");
    size_t k = 0;
    for (int m = 0; m < nmembers && !rc; ++m) {
        BatchMember& bm = B->h_members[m];
        char* base = (char*)B->d_scratch;
        bm.xT = (float*)(base + off_xT[m]);
        bm.pf = (double*)(base + off_pf[m]);
        bm.pb = (double*)(base + off_pb[m]);
        for (int z = 0; z < bm.nslices; ++z, ++k) {
            refs[k].pf = bm.pf + (size_t)3 * bm.bps_f * z;
            refs[k].pb = bm.pb + (size_t)3 * bm.bps_b * z;
        }
    }
    if (!rc) rc = check_hip(hipMalloc(&B->d_members, sizeof(BatchMember) * nmembers), "batch_create");
    if (!rc) rc = check_hip(hipMalloc(&B->d_fwd, sizeof(BatchItem) * fwd.size()), "batch_create");
    if (!rc) rc = check_hip(hipMalloc(&B->d_back, sizeof(BatchItem) * back.size()), "batch_create");
    if (!rc) rc = check_hip(hipMalloc(&B->d_slices, sizeof(SliceRef) * refs.size()), "batch_create");
    if (!rc) rc = check_hip(hipMalloc(&B->d_counter, 2 * sizeof(int)), "batch_create");
    if (!rc) rc = check_hip(hipMemcpy(B->d_members, B->h_members.data(), sizeof(BatchMember) * nmembers, hipMemcpyHostToDevice), "batch_create");
    if (!rc) rc = check_hip(hipMemcpy(B->d_fwd, fwd.data(), sizeof(BatchItem) * fwd.size(), hipMemcpyHostToDevice), "batch_create");
    if (!rc) rc = check_hip(hipMemcpy(B->d_back, back.data(), sizeof(BatchItem) * back.size(), hipMemcpyHostToDevice), "batch_create");
    if (!rc) rc = check_hip(hipMemcpy(B->d_slices, refs.data(), sizeof(SliceRef) * refs.size(), hipMemcpyHostToDevice), "batch_create");
    int dev = 0, cus = 256;
    if (!rc) rc = check_hip(hipGetDevice(&dev), "batch_create");
    if (!rc) rc = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "batch_create");
    B->grid = cus * 8;
    if (rc) {
        itomo_batch_destroy(B);
        return rc;
    }
    *out = B;
    return ITOMO_OK;
}

int itomo_batch_destroy(itomo_batch* B)
{
    if (!B) return ITOMO_OK;
    (void)hipFree(B->d_members);
    (void)hipFree(B->d_fwd);
    (void)hipFree(B->d_back);
    (void)hipFree(B->d_slices);
    (void)hipFree(B->d_counter);
    (void)hipFree(B->d_scratch);
    delete B;
    return ITOMO_OK;
}

int itomo_batch_sirt_iterate(itomo_batch* B, float pc_beta, void* stream)
{
    if (!B) return set_error(ITOMO_EINVAL, "batch_sirt_iterate: null batch");
    hipStream_t st = (hipStream_t)stream;
    for (int m = 0; m < B->nmembers; ++m) {
        const BatchMember& bm = B->h_members[m];
        dim3 tb(32, 8), tg((bm.npix + 31) / 32, (bm.npix + 31) / 32, bm.nslices);
        transpose_kernel_b<<<tg, tb, 0, st>>>(bm.x, bm.xT, bm.npix, (int64_t)bm.npix * bm.npix);
    }
    reset_counter_kernel<<<1, 1, 0, st>>>(B->d_counter);
    batch_fwd_kernel<<<std::min(B->grid * (256 / kBatchFwdBlock), B->nfwd), kBatchFwdBlock, 0,
                       st>>>(B->d_members, B->d_fwd, B->nfwd, B->d_counter);
    batch_back_kernel<<<std::min(B->grid, B->nback), 256, 0, st>>>(B->d_members, B->d_back,
                                                                     B->nback, B->d_counter,
                                                                     pc_beta);
    batch_finalize_kernel<<<B->total_slices, 256, 0, st>>>(B->d_slices);
    return check_hip(hipGetLastError(), "batch_sirt_iterate launch");
}

// ---- hipGraph capture -------------------------------------------------------------------

int itomo_sirt_iterate(const itomo_plan*, float*, const float*, const float*, const float*,
                       float*, int, float, double*, void*);

// Captures `iters` calls of enqueue(ctx) into a graph.  `scratch` (may be null) is a
// device region owned by the new graph: freed by itomo_graph_destroy, or here on failure.
static int capture(itomo_graph** out, void* stream, int iters,
                   int (*enqueue)(void* ctx, hipStream_t st), void* ctx, void* scratch)
{
    if (!out || iters < 1) {
        if (scratch) (void)hipFree(scratch);
        return set_error(ITOMO_EINVAL, "graph: bad argument");
    }
    *out = nullptr;
    hipStream_t st = (hipStream_t)stream, cap = nullptr;
    int rc = check_hip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "graph: stream");
    if (rc) return rc;
    // start the capture after the caller's queued work
    hipEvent_t ev = nullptr;
    rc = check_hip(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "graph: event");
    if (!rc) rc = check_hip(hipEventRecord(ev, st), "graph: record");
    if (!rc) rc = check_hip(hipStreamWaitEvent(cap, ev, 0), "graph: wait");
    if (!rc) rc = check_hip(hipStreamSynchronize(cap), "graph: sync");
    itomo_graph* G = new (std::nothrow) itomo_graph();
    if (!G) rc = set_error(ITOMO_ENOMEM, "graph: out of host memory");
    if (!rc) rc = check_hip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "graph: begin");
    int rc_enq = ITOMO_OK;
    if (!rc) {
        for (int i = 0; i < iters && !rc_enq; ++i) rc_enq = enqueue(ctx, cap);
        hipGraph_t g = nullptr;
        int rc_end = check_hip(hipStreamEndCapture(cap, &g), "graph: end");
        G->graph = g;
        rc = rc_enq ? rc_enq : rc_end;
    }
    if (!rc) rc = check_hip(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0), "graph: instantiate");
    if (ev) (void)hipEventDestroy(ev);
    if (cap) (void)hipStreamDestroy(cap);
    if (rc) {
        if (G) {
            if (G->exec) (void)hipGraphExecDestroy(G->exec);
            if (G->graph) (void)hipGraphDestroy(G->graph);
            delete G;
        }
        if (scratch) (void)hipFree(scratch);
        return rc;
    }
    G->scratch = scratch;
    *out = G;
    return ITOMO_OK;
}

struct SirtCtx {
    const itomo_plan* p;
    float* x;
    const float* b;
    const float* R;
    const float* C;
    float* work;
    int nslices;
    float pc;
    double* stats;
};

static int enqueue_sirt(void* c, hipStream_t st)
{
    SirtCtx* s = (SirtCtx*)c;
    return itomo_sirt_iterate(s->p, s->x, s->b, s->R, s->C, s->work, s->nslices, s->pc,
                              s->stats, st);
}

int itomo_graph_sirt(itomo_graph** out, const itomo_plan* p, float* x, const float* b,
                     const float* R, const float* C, float* work, int nslices, float pc_beta,
                     double* stats, int iters, void* stream)
{
    if (!p) return set_error(ITOMO_EINVAL, "graph_sirt: null plan");
    if (nslices < 1 || nslices > 65535)
        return set_error(ITOMO_EINVAL, "graph_sirt: nslices %d outside [1, 65535]", nslices);
    // The graph gets its own scratch, allocated before capture (nothing is allocated
    // inside a captured region) and freed with the graph: later calls on the plan may
    // grow or re-lay out the plan's own buffer without touching what the graph replays.
    const size_t bytes = scratch_total(*p, nslices);
    void* scratch = nullptr;
    int rc = check_hip(hipMalloc(&scratch, bytes), "graph_sirt: hipMalloc scratch");
    if (rc) return rc;
    ScratchScope scope(ScratchRegion{scratch, bytes});
    SirtCtx ctx{p, x, b, R, C, work, nslices, pc_beta, stats};
    return capture(out, stream, iters, enqueue_sirt, &ctx, scratch);
}

struct BatchCtx {
    itomo_batch* B;
    float pc;
};

static int enqueue_batch(void* c, hipStream_t st)
{
    BatchCtx* s = (BatchCtx*)c;
    return itomo_batch_sirt_iterate(s->B, s->pc, st);
}

int itomo_graph_batch_sirt(itomo_graph** out, itomo_batch* B, float pc_beta, int iters,
                           void* stream)
{
    if (!B) return set_error(ITOMO_EINVAL, "graph_batch_sirt: null batch");
    BatchCtx ctx{B, pc_beta};
    return capture(out, stream, iters, enqueue_batch, &ctx, nullptr);
}

int itomo_graph_launch(itomo_graph* G, void* stream)
{
    if (!G || !G->exec) return set_error(ITOMO_EINVAL, "graph_launch: null graph");
    return check_hip(hipGraphLaunch(G->exec, (hipStream_t)stream), "graph_launch");
}

int itomo_graph_destroy(itomo_graph* G)
{
    if (!G) return ITOMO_OK;
    if (G->exec) (void)hipGraphExecDestroy(G->exec);
    if (G->graph) (void)hipGraphDestroy(G->graph);
    if (G->scratch) (void)hipFree(G->scratch);
    delete G;
    return ITOMO_OK;
}

}  // extern "C"
